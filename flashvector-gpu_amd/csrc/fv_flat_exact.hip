// Exact FP32 flat scan fused with top-k (the cuda_flat_scan_batch path).
//
// Reference: kernels/src/sparse_hybrid.cu:13-107 (one block per query, one
// vector per warp, thread-0 serial insertion). Here the scan is a tiled
// queries x rows sweep: a workgroup owns 64 queries and one chunk of rows,
// every thread holds a 4x4 register tile of exact distances accumulated in the
// reference CPU path's order (sequential over d, non-fused), and a candidate
// that beats its query's running threshold is appended to a per-(chunk, query)
// buffer. Tiles are staged through LDS 32 dims at a time; the next stage's
// rows and queries are loaded into registers (128-bit loads) while the
// current stage is computed. Buffers that could overflow are cut by one
// wavefront (bisection over the distance, exact sort on heavy ties). A second
// kernel merges the per-chunk lists of each query.
#include <mutex>

#include "fv_flat_exact.hpp"
#include "fv_workspace.hpp"

namespace fv {

constexpr int kTQ = 64;        // queries per workgroup
constexpr int kTN = 64;        // rows per tile
constexpr int kDK = 32;        // dims per LDS stage
constexpr int kLdsStride = 68; // padded LDS row (floats), keeps float4 alignment
constexpr int kScanThreads = 256;
// Static LDS of the scan kernel: staged queries and rows, counts, thresholds.
constexpr uint32_t kExactStageFloats = 2u * kDK * kLdsStride;
constexpr uint32_t kExactStaticLds = kExactStageFloats * 4u + kTQ * 8u;
// Largest buffer capacity whose 4 wave-private sort scratches fit in the
// staging area (they are used only between tiles, when it is idle).
constexpr uint32_t kExactAliasCap = kExactStageFloats * 4u / (4u * 8u);
constexpr uint32_t exact_dynamic_lds(uint32_t cap) { return cap <= kExactAliasCap ? 0u : 4u * cap * 8u; }
constexpr uint32_t kCUs = 256;           // MI355X compute units
constexpr uint32_t kLdsPerCU = 160u * 1024u;

// Row norms (cosine): sequential non-fused sums, bit-identical to the sums the
// oracle forms inside its per-pair loop.
__global__ void row_norms_kernel(const float* __restrict__ x, uint32_t n, uint32_t dim,
                                 float* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = x + (size_t)i * dim;
    float acc = 0.0f;
    for (uint32_t d = 0; d < dim; ++d) {
        float v = r[d];
        float sq = v * v;
        acc = acc + sq;
    }
    out[i] = acc;
}

// One stage (64 rows and 64 queries x kDK dims) held in registers: thread t
// holds dims 4*(t % kSegs) .. +3 of tile rows t / kSegs + i * kRowsPerPass,
// zero past the row range or the dimension.
constexpr int kSegs = kDK / 4;
constexpr int kRowsPerPass = kScanThreads / kSegs;
constexpr int kPasses = kTN / kRowsPerPass;
static_assert(kTN == kTQ && kPasses * kRowsPerPass == kTN, "stage shape");
struct ExactStage {
    float x[kPasses][4];
    float q[kPasses][4];
};

__device__ __forceinline__ void load_exact_stage(const float* __restrict__ X, const float* __restrict__ Qm,
                                                 uint32_t dim, uint32_t nq, uint32_t q0, uint32_t r0,
                                                 uint32_t r_end, uint32_t d0, bool vec, int tid,
                                                 ExactStage& s) {
    const uint32_t dg = d0 + (uint32_t)(tid % kSegs) * 4u;
#pragma unroll
    for (int it = 0; it < kPasses; ++it) {
        const uint32_t rr = (uint32_t)(tid / kSegs) + (uint32_t)(it * kRowsPerPass);
        const uint32_t rg = r0 + rr, qg = q0 + rr;
        if (vec) {
            // dim is a multiple of 4, so a vector that starts in range ends in range.
            float4 xv = make_float4(0.0f, 0.0f, 0.0f, 0.0f), qv = xv;
            if (rg < r_end && dg < dim) xv = *reinterpret_cast<const float4*>(X + (size_t)rg * dim + dg);
            if (qg < nq && dg < dim) qv = *reinterpret_cast<const float4*>(Qm + (size_t)qg * dim + dg);
            s.x[it][0] = xv.x; s.x[it][1] = xv.y; s.x[it][2] = xv.z; s.x[it][3] = xv.w;
            s.q[it][0] = qv.x; s.q[it][1] = qv.y; s.q[it][2] = qv.z; s.q[it][3] = qv.w;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                s.x[it][c] = (rg < r_end && dg + c < dim) ? X[(size_t)rg * dim + dg + c] : 0.0f;
                s.q[it][c] = (qg < nq && dg + c < dim) ? Qm[(size_t)qg * dim + dg + c] : 0.0f;
            }
        }
    }
}

// Transposed into LDS: sx[d][row], sq[d][query].
__device__ __forceinline__ void store_exact_stage(float* sq, float* sx, int tid, const ExactStage& s) {
    const int dd = (tid % kSegs) * 4;
#pragma unroll
    for (int it = 0; it < kPasses; ++it) {
        const int rr = tid / kSegs + it * kRowsPerPass;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            sx[(dd + c) * kLdsStride + rr] = s.x[it][c];
            sq[(dd + c) * kLdsStride + rr] = s.q[it][c];
        }
    }
}

// Cuts one (chunk, query) buffer of c keys (c <= cap, cap >= keep + 128) so
// that another tile of appends fits, by one wavefront: bisection keeps at
// most keep + 64 keys at or below a bound; when ties at the bound keep more,
// an exact sort keeps `keep`. *bound receives an ord(distance) that no later
// row of the chunk may reach (0xFFFFFFFF: none yet).
__device__ __forceinline__ uint32_t prune_exact_buffer(uint64_t* buf, uint32_t c, uint32_t keep,
                                                       uint32_t cap, uint64_t* scratch, int lane,
                                                       uint32_t* bound) {
    uint32_t cnt = c;
    *bound = 0xFFFFFFFFu;
    if (cap <= 512u) {
        cnt = cap <= 256u ? wave_bisect_prune<4>(buf, c, keep, lane, bound)
                          : wave_bisect_prune<8>(buf, c, keep, lane, bound);
        if (cnt + kTN <= cap) return cnt;
        __threadfence_block();  // the compacted keys are read back by other lanes
    }
    const uint32_t kept = wave_prune(buf, cnt, keep, scratch, lane);
    *bound = kept == keep ? (uint32_t)(scratch[keep - 1] >> 32) : 0xFFFFFFFFu;
    return kept;
}

// FMA = true accumulates L2 as fmaf(diff, diff, acc), the order of the IVF
// coarse quantizer (kernels/src/ivf_rerank.cuh:81-84); false is the reference
// CPU path's non-fused order.
template <int METRIC, bool FMA>
__global__ __launch_bounds__(kScanThreads) void flat_exact_scan_kernel(
    const float* __restrict__ X, uint32_t n, uint32_t dim, const float* __restrict__ Qm,
    uint32_t nq, const float* __restrict__ xnorm, const float* __restrict__ qnorm,
    const uint32_t* __restrict__ mask, uint32_t keep, uint32_t cap, uint32_t rows_per_chunk,
    uint64_t* __restrict__ cand, uint64_t* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) float s_stage[kExactStageFloats];
    __shared__ uint32_t s_cnt[kTQ];
    __shared__ float s_thr[kTQ];
    extern __shared__ uint64_t s_scratch[];  // 4 waves x cap keys when cap > kExactAliasCap
    float* sq = s_stage;
    float* sx = s_stage + kDK * kLdsStride;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const uint32_t q0 = blockIdx.x * kTQ;
    const uint32_t chunk = blockIdx.y;
    const uint32_t r_begin = chunk * rows_per_chunk;
    const uint32_t r_end = min(n, r_begin + rows_per_chunk);
    const int tq = (tid & 15) * 4;
    const int tr = (tid >> 4) * 4;
    // Sort scratch: the staging area between tiles (every wave has passed the
    // barrier after the tile's last reads, and the next tile is stored after
    // the next barrier), or dynamic LDS for large capacities.
    uint64_t* my_scratch = (cap <= kExactAliasCap ? reinterpret_cast<uint64_t*>(s_stage) : s_scratch) +
                           (size_t)wave * cap;
    const bool vec = (dim & 3u) == 0u && (reinterpret_cast<uintptr_t>(X) & 15u) == 0u &&
                     (reinterpret_cast<uintptr_t>(Qm) & 15u) == 0u;

    if (tid < kTQ) {
        s_cnt[tid] = 0;
        s_thr[tid] = kAdmit;
    }
    float qn[4];
    if (METRIC == METRIC_COSINE) {
        for (int i = 0; i < 4; ++i) qn[i] = (q0 + tq + i < nq) ? qnorm[q0 + tq + i] : 1.0f;
    }
    ExactStage st;
    if (r_begin < r_end) load_exact_stage(X, Qm, dim, nq, q0, r_begin, r_end, 0u, vec, tid, st);
    for (uint32_t r0 = r_begin; r0 < r_end; r0 += kTN) {
        float acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
        for (uint32_t d0 = 0; d0 < dim; d0 += kDK) {
            __syncthreads();
            store_exact_stage(sq, sx, tid, st);
            __syncthreads();
            // Next stage: the next dims of this tile, else the next tile's first.
            uint32_t nd0 = d0 + kDK, nr0 = r0;
            if (nd0 >= dim) {
                nd0 = 0;
                nr0 = r0 + kTN;
            }
            if (nr0 < r_end) load_exact_stage(X, Qm, dim, nq, q0, nr0, r_end, nd0, vec, tid, st);
            const int dlim = min((int)kDK, (int)(dim - d0));
            for (int dd = 0; dd < dlim; ++dd) {
                float4 qv = *reinterpret_cast<const float4*>(&sq[dd * kLdsStride + tq]);
                float4 xv = *reinterpret_cast<const float4*>(&sx[dd * kLdsStride + tr]);
                float qa[4] = {qv.x, qv.y, qv.z, qv.w};
                float xa[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (METRIC == METRIC_L2 && FMA) {
                            float diff = qa[i] - xa[j];
                            acc[i][j] = __builtin_fmaf(diff, diff, acc[i][j]);
                        } else if (METRIC == METRIC_L2) {
                            float diff = qa[i] - xa[j];
                            float s = diff * diff;
                            acc[i][j] = acc[i][j] + s;
                        } else {
                            float p = qa[i] * xa[j];
                            acc[i][j] = acc[i][j] + p;
                        }
                    }
            }
        }
        // Filter against each query's threshold and append survivors.
        bool row_ok[4];
        float xn[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t r = r0 + tr + j;
            row_ok[j] = r < r_end && mask_allows(mask, r);
            if (METRIC == METRIC_COSINE) xn[j] = row_ok[j] ? xnorm[r] : 1.0f;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint32_t qg = q0 + tq + i;
            if (qg >= nq) continue;
            float thr = s_thr[tq + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float dist;
                if (METRIC == METRIC_L2) dist = acc[i][j];
                else if (METRIC == METRIC_INNER_PRODUCT) dist = -acc[i][j];
                else dist = 1.0f - acc[i][j] / (sqrtf(qn[i]) * sqrtf(xn[j]) + 1e-8f);
                if (row_ok[j] && dist < thr) {
                    uint32_t pos = atomicAdd(&s_cnt[tq + i], 1u);
                    cand[((size_t)chunk * nq + qg) * cap + pos] = make_key(dist, r0 + tr + j);
                }
            }
        }
        __syncthreads();
        // Cut every buffer that could overflow during the next tile.
        for (int qq = wave; qq < kTQ; qq += 4) {
            const uint32_t qg = q0 + qq;
            if (qg >= nq) break;
            const uint32_t c = s_cnt[qq];
            if (c + kTN <= cap) continue;
            uint64_t* buf = cand + ((size_t)chunk * nq + qg) * cap;
            uint32_t t;
            const uint32_t kept = prune_exact_buffer(buf, c, keep, cap, my_scratch, lane, &t);
            if (lane == 0) {
                s_cnt[qq] = kept;
                if (t != 0xFFFFFFFFu) s_thr[qq] = fminf(s_thr[qq], unord_f32(t));
            }
            wave_lds_sync();
        }
    }
    __syncthreads();
    // Emit each query's sorted `keep` best of this chunk: bisection first, so
    // that at most keep + 64 keys (or the tied ones) are sorted.
    for (int qq = wave; qq < kTQ; qq += 4) {
        uint32_t qg = q0 + qq;
        if (qg >= nq) break;
        uint32_t c = s_cnt[qq];
        uint64_t* buf = cand + ((size_t)chunk * nq + qg) * cap;
        if (c > keep && cap <= 512u) {
            uint32_t t;
            c = cap <= 256u ? wave_bisect_prune<4>(buf, c, keep, lane, &t)
                            : wave_bisect_prune<8>(buf, c, keep, lane, &t);
            __threadfence_block();
        }
        uint32_t kept = wave_prune(buf, c, keep, my_scratch, lane);
        uint64_t* out = partial + ((size_t)chunk * nq + qg) * keep;
        for (uint32_t i = lane; i < keep; i += kWave) out[i] = i < kept ? my_scratch[i] : kEmptyKey;
    }
}

// Merges the per-chunk sorted lists of one query (one workgroup per query) and
// writes the first top_k results, remapping slots through vec_ids.
__global__ __launch_bounds__(256) void merge_partials_kernel(
    const uint64_t* __restrict__ partial, uint32_t n_chunks, uint32_t nq, uint32_t keep,
    uint32_t top_k, uint32_t out_stride, const uint32_t* __restrict__ vec_ids,
    GpuSearchResult* __restrict__ out) {
    extern __shared__ uint64_t s_keys[];
    const uint32_t q = blockIdx.x;
    const uint32_t total = n_chunks * keep;
    const uint32_t n = next_pow2(total);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        uint64_t k = kEmptyKey;
        if (i < total) {
            uint32_t c = i / keep, j = i % keep;
            k = partial[((size_t)c * nq + q) * keep + j];
        }
        s_keys[i] = k;
    }
    __syncthreads();
    block_bitonic_sort(s_keys, n);
    for (uint32_t i = threadIdx.x; i < top_k; i += blockDim.x) {
        uint64_t k = s_keys[i];
        GpuSearchResult r;
        if (k == kEmptyKey) {
            r.id = FV_EMPTY_ID;
            r.distance = FV_EMPTY_DIST;
        } else {
            uint32_t slot = key_slot(k);
            r.id = vec_ids ? vec_ids[slot] : slot;
            r.distance = key_dist(k);
        }
        out[(size_t)q * out_stride + i] = r;
    }
}

FlatExactPlan plan_flat_exact(uint32_t n, uint32_t dim, uint32_t nq, uint32_t top_k, int metric,
                              size_t max_bytes) {
    FlatExactPlan p{};
    p.keep = top_k;
    // Bisection prunes leave up to keep + 64 keys; one more tile adds 64.
    p.cap = next_pow2(top_k + 2u * kTN);
    uint32_t qblocks = (nq + kTQ - 1) / kTQ;
    // About 2048 workgroups, rounded to whole rounds of the workgroups the
    // CUs hold at once (LDS-bound; at most 8 of 4 wavefronts per CU), so that
    // the last round is not a sparse one.
    const uint32_t lds = kExactStaticLds + exact_dynamic_lds(p.cap);
    uint32_t per_cu = kLdsPerCU / lds;
    if (per_cu > 8u) per_cu = 8u;
    if (per_cu < 1u) per_cu = 1u;
    const uint32_t slots = kCUs * per_cu;
    const uint32_t want_total = qblocks * ((2048u + qblocks - 1) / qblocks);
    uint32_t rounds = (want_total + slots / 2) / slots;
    if (rounds < 1u) rounds = 1u;
    uint32_t want = rounds * slots / qblocks;
    uint32_t max_by_rows = (n + 255u) / 256u;
    uint32_t max_by_merge = 8192u / top_k;
    uint32_t chunks = want;
    if (chunks > max_by_rows) chunks = max_by_rows;
    if (chunks > max_by_merge) chunks = max_by_merge;
    if (chunks < 1) chunks = 1;
    // Workspace cap: candidate buffers and partial lists are per (chunk,
    // query), the cosine norms fixed.
    const size_t fixed = metric == METRIC_COSINE ? ((size_t)n + nq) * sizeof(float) + 256 : 256;
    const size_t per_chunk = (size_t)nq * (p.cap + p.keep) * sizeof(uint64_t);
    if (max_bytes != ~(size_t)0) {
        const size_t fit = max_bytes > fixed ? (max_bytes - fixed) / per_chunk : 0;
        if (fit < 1) {
            p.workspace_bytes = 0;
            return p;
        }
        if (chunks > fit) chunks = (uint32_t)fit;
    }
    uint32_t rows = (n + chunks - 1) / chunks;
    rows = (rows + kTN - 1) / kTN * kTN;
    p.rows_per_chunk = rows;
    p.n_chunks = (n + rows - 1) / rows;
    p.qblocks = qblocks;
    size_t off = 0;
    p.off_cand = off;
    off += (size_t)p.n_chunks * nq * p.cap * sizeof(uint64_t);
    p.off_partial = off;
    off += (size_t)p.n_chunks * nq * p.keep * sizeof(uint64_t);
    p.off_norms = off;
    if (metric == METRIC_COSINE) off += ((size_t)n + nq) * sizeof(float);
    off = (off + 255) / 256 * 256;
    p.workspace_bytes = off;
    (void)dim;
    return p;
}

hipError_t launch_flat_exact(const float* X, const uint32_t* vec_ids, uint32_t n, uint32_t dim,
                             const float* Qm, uint32_t nq, uint32_t top_k, int metric,
                             const uint32_t* mask, GpuSearchResult* out, uint32_t out_stride,
                             void* workspace, const FlatExactPlan& p, hipStream_t stream,
                             bool fma_l2) {
    char* ws = static_cast<char*>(workspace);
    uint64_t* cand = reinterpret_cast<uint64_t*>(ws + p.off_cand);
    uint64_t* partial = reinterpret_cast<uint64_t*>(ws + p.off_partial);
    float* xnorm = nullptr;
    float* qnorm = nullptr;
    if (metric == METRIC_COSINE) {
        xnorm = reinterpret_cast<float*>(ws + p.off_norms);
        qnorm = xnorm + n;
        row_norms_kernel<<<(n + 255) / 256, 256, 0, stream>>>(X, n, dim, xnorm);
        row_norms_kernel<<<(nq + 255) / 256, 256, 0, stream>>>(Qm, nq, dim, qnorm);
    }
    dim3 grid(p.qblocks, p.n_chunks);
    const size_t lds = exact_dynamic_lds(p.cap);
    void* timer = fma_l2 ? nullptr : timer_begin(stream);
    if (metric == METRIC_INNER_PRODUCT)
        flat_exact_scan_kernel<METRIC_INNER_PRODUCT, false><<<grid, kScanThreads, lds, stream>>>(
            X, n, dim, Qm, nq, xnorm, qnorm, mask, p.keep, p.cap, p.rows_per_chunk, cand, partial);
    else if (metric == METRIC_COSINE)
        flat_exact_scan_kernel<METRIC_COSINE, false><<<grid, kScanThreads, lds, stream>>>(
            X, n, dim, Qm, nq, xnorm, qnorm, mask, p.keep, p.cap, p.rows_per_chunk, cand, partial);
    else if (fma_l2)
        flat_exact_scan_kernel<METRIC_L2, true><<<grid, kScanThreads, lds, stream>>>(
            X, n, dim, Qm, nq, xnorm, qnorm, mask, p.keep, p.cap, p.rows_per_chunk, cand, partial);
    else
        flat_exact_scan_kernel<METRIC_L2, false><<<grid, kScanThreads, lds, stream>>>(
            X, n, dim, Qm, nq, xnorm, qnorm, mask, p.keep, p.cap, p.rows_per_chunk, cand, partial);
    timer_end(timer, stream);
    uint32_t total = next_pow2(p.n_chunks * p.keep);
    merge_partials_kernel<<<nq, 256, (size_t)total * sizeof(uint64_t), stream>>>(
        partial, p.n_chunks, nq, p.keep, top_k, out_stride, vec_ids, out);
    return hipGetLastError();
}

// ---- Static scratch of fv_flat_scan_batch_noalloc --------------------------
// A file-scope device array, one copy per GPU, held until the process ends.
// Its host-side name is only a stand-in, never a device address: the address
// is taken with hipGetSymbolAddress once per device (with that device
// current) and only that pointer reaches kernels. Nothing in it needs to
// start at zero: the scan writes every candidate before reading it (counts
// live in LDS) and writes every partial-list slot.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
constexpr int kNoAllocMaxDevices = 64;
static std::once_flag g_scratch_once[kNoAllocMaxDevices];
static void* g_scratch_ptr[kNoAllocMaxDevices];
static hipError_t g_scratch_err[kNoAllocMaxDevices];

static hipError_t noalloc_scratch(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::call_once(g_scratch_once[dev], [dev] {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? p : nullptr;
    });
    *out = g_scratch_ptr[dev];
    return g_scratch_err[dev];
}

}  // namespace fv

extern "C" {

// No counterpart in the reference: cuda_flat_scan_batch for callers that
// cannot allocate device memory (HIP graph capture, serving loops that forbid
// allocation). Same arguments, clamps and results; its scratch is the
// library's static 64 MiB per-GPU array, so it calls no allocator. Calls that
// overlap on one GPU share that array and race: serialize them (one stream,
// or host-side ordering). A call whose scratch cannot fit even one row chunk
// (about nq * (next_pow2(k + 128) + k) * 8 bytes, plus (n + nq) * 4 for
// cosine) returns hipErrorOutOfMemory and launches nothing.
int fv_flat_scan_batch_noalloc(const float* d_vectors, const uint32_t* d_vec_ids,
                               uint32_t num_vectors, uint32_t dim, const float* d_queries,
                               uint32_t num_queries, uint32_t top_k, MetricType metric,
                               const uint32_t* d_valid_mask, GpuSearchResult* d_out_results,
                               void* stream) {
    if (!d_vectors || !d_queries || !d_out_results) return -1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_queries == 0 || top_k == 0) return 0;
    if (top_k > 256u) top_k = 256u;
    if (num_vectors == 0) {
        hipError_t e = fv::fill_results(d_out_results, (size_t)num_queries * top_k, 0xFFFFFFFFu,
                                        0xFFFFFFFFu, s);
        return e == hipSuccess ? 0 : (int)e;
    }
    if (dim == 0) return (int)hipErrorInvalidValue;
    fv::FlatExactPlan p = fv::plan_flat_exact(num_vectors, dim, num_queries, top_k, (int)metric,
                                              fv::kNoAllocScratchBytes);
    if (p.workspace_bytes == 0 || p.workspace_bytes > fv::kNoAllocScratchBytes)
        return (int)hipErrorOutOfMemory;
    void* ws = nullptr;
    hipError_t e = fv::noalloc_scratch(&ws);
    if (e != hipSuccess) return (int)e;
    e = fv::launch_flat_exact(d_vectors, d_vec_ids, num_vectors, dim, d_queries, num_queries,
                              top_k, (int)metric, d_valid_mask, d_out_results, top_k, ws, p, s,
                              false);
    return e == hipSuccess ? 0 : (int)e;
}

// cuda_bridge.h:150-162 / kernels/src/sparse_hybrid.cu:504-524.
int cuda_flat_scan_batch(const float* d_vectors, const uint32_t* d_vec_ids, uint32_t num_vectors,
                         uint32_t dim, const float* d_queries, uint32_t num_queries,
                         uint32_t top_k, MetricType metric, const uint32_t* d_valid_mask,
                         GpuSearchResult* d_out_results, void* stream) {
    if (!d_vectors || !d_queries || !d_out_results) return -1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_queries == 0 || top_k == 0) return 0;
    if (top_k > 256u) top_k = 256u;
    if (num_vectors == 0) {
        // All bytes 0xFF, as the reference's cudaMemsetAsync (sparse_hybrid.cu:98-101);
        // a fill kernel rather than a memset node keeps the call graph-capturable.
        hipError_t e = fv::fill_results(d_out_results, (size_t)num_queries * top_k, 0xFFFFFFFFu,
                                        0xFFFFFFFFu, s);
        return e == hipSuccess ? 0 : (int)e;
    }
    if (dim == 0) return (int)hipErrorInvalidValue;
    fv::FlatExactPlan p = fv::plan_flat_exact(num_vectors, dim, num_queries, top_k, (int)metric);
    fv::WorkspaceLease ws;
    hipError_t e = fv::stream_workspace(s, p.workspace_bytes, ws);
    if (e != hipSuccess) return (int)e;
    e = fv::launch_flat_exact(d_vectors, d_vec_ids, num_vectors, dim, d_queries, num_queries,
                              top_k, (int)metric, d_valid_mask, d_out_results, top_k, ws.ptr, p, s,
                              false);
    return e == hipSuccess ? 0 : (int)e;
}

size_t fv_flat_exact_workspace_size(uint32_t num_vectors, uint32_t dim, uint32_t num_queries,
                                    uint32_t top_k, MetricType metric) {
    if (top_k == 0 || top_k > FV_MAX_TOPK || num_vectors == 0) return 0;
    return fv::plan_flat_exact(num_vectors, dim, num_queries, top_k, (int)metric).workspace_bytes;
}

int fv_flat_exact_search(const float* d_vectors, uint32_t num_vectors, uint32_t dim,
                         const float* d_queries, uint32_t num_queries, uint32_t top_k,
                         MetricType metric, const uint32_t* d_valid_mask,
                         GpuSearchResult* d_out_results, void* d_workspace,
                         size_t workspace_bytes, void* stream) {
    if (!d_vectors || !d_queries || !d_out_results) return -1;
    if (top_k > FV_MAX_TOPK || dim == 0) return (int)hipErrorInvalidValue;
    if (num_queries == 0 || top_k == 0) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_vectors == 0) {
        // All bytes 0xFF, as the reference's cudaMemsetAsync (sparse_hybrid.cu:98-101);
        // a fill kernel rather than a memset node keeps the call graph-capturable.
        hipError_t e = fv::fill_results(d_out_results, (size_t)num_queries * top_k, 0xFFFFFFFFu,
                                        0xFFFFFFFFu, s);
        return e == hipSuccess ? 0 : (int)e;
    }
    fv::FlatExactPlan p = fv::plan_flat_exact(num_vectors, dim, num_queries, top_k, (int)metric);
    if (!d_workspace || workspace_bytes < p.workspace_bytes) return (int)hipErrorInvalidValue;
    hipError_t e = fv::launch_flat_exact(d_vectors, nullptr, num_vectors, dim, d_queries,
                                         num_queries, top_k, (int)metric, d_valid_mask,
                                         d_out_results, top_k, d_workspace, p, s, false);
    return e == hipSuccess ? 0 : (int)e;
}

}  // extern "C"
