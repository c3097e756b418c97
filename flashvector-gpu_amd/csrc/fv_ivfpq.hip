// IVF-PQ ADC list scan fused with shortlist selection and exact FP32 rerank
// (cuda_ivf_pq_search_batch and fv_ivf_pq_search).
//
// Reference: kernels/src/ivf_pq_lookup.cu:14-475 with the coarse quantizer of
// kernels/src/ivf_rerank.cuh:49-130. The reference keeps per-thread candidate
// lists merged serially (shortlist <= 32) or appends under a racy cutoff with
// global-memory bitonic prunes, and reads each entry's codes by vector id.
//
// Here, per query:
//  * ivf_coarse_kernel: every centroid distance as an fmaf chain over d (the
//    reference's order), then the nprobe smallest by (distance, cluster) from
//    one LDS bitonic sort.
//  * ivf_scan_kernel (one workgroup per query; the hardware dispatcher hands
//    queries of uneven cost to free CUs as they drain):
//      for each probed list: the residual LUT [M][256] is built in LDS (fmaf
//      chain over the sub-dimension, the reference's order), then the list is
//      streamed 256 entries per pass. With a list-ordered code copy every
//      lane reads its entry's 16 code bytes with one coalesced 128-bit load;
//      with the reference's id-indexed codes it gathers them by vector id.
//      ADC = sequential FP32 adds over m = 0..M-1. An entry is kept if its
//      (adc, id) key is below the running threshold key; keys go to a
//      candidate buffer (LDS, or global memory for very deep shortlists) that
//      the workgroup sorts and truncates to S when it nears capacity.
//    finally the S best are reranked with the exact FP32 distance (the
//    reference CPU path's order), sorted by (distance, id) and the first k
//    are written.
// No MFMA: the scan is bandwidth- and LDS-lookup-bound.
#include "fv_ivfpq.hpp"

namespace fv {

constexpr int kIvfThreads = 256;
constexpr uint32_t kPassesPerCheck = 4;  // 256-entry passes between capacity checks
constexpr uint32_t kEntriesPerCheck = kIvfThreads * kPassesPerCheck;

__global__ __launch_bounds__(256) void ivf_coarse_kernel(
    const float* __restrict__ C, uint32_t nlist, uint32_t dim, const float* __restrict__ Qm,
    const uint32_t* __restrict__ nprobe_q, uint32_t nprobe_u, uint32_t max_nprobe, uint32_t npow,
    uint32_t* __restrict__ probes) {
    extern __shared__ uint64_t s_ck[];
    const uint32_t q = blockIdx.x;
    const float* qv = Qm + (size_t)q * dim;
    for (uint32_t c = threadIdx.x; c < npow; c += blockDim.x) {
        uint64_t key = kEmptyKey;
        if (c < nlist) {
            const float* cen = C + (size_t)c * dim;
            float dist = 0.0f;
            for (uint32_t d = 0; d < dim; ++d) {
                float diff = qv[d] - cen[d];
                dist = __builtin_fmaf(diff, diff, dist);
            }
            if (dist < kAdmit) key = make_key(dist, c);
        }
        s_ck[c] = key;
    }
    __syncthreads();
    block_bitonic_sort(s_ck, npow);
    uint32_t np = nprobe_q ? nprobe_q[q] : nprobe_u;
    if (np == 0) np = 1;
    if (np > nlist) np = nlist;
    for (uint32_t i = threadIdx.x; i < max_nprobe; i += blockDim.x) {
        uint64_t k = i < np ? s_ck[i] : kEmptyKey;
        probes[(size_t)q * max_nprobe + i] = k == kEmptyKey ? FV_EMPTY_ID : key_slot(k);
    }
}

// Sorts the buffer, keeps the `keep` smallest and returns the new threshold
// key (kEmptyKey while fewer than `keep` candidates exist). All threads call.
__device__ __forceinline__ uint64_t block_prune(uint64_t* buf, uint32_t* s_cnt, uint32_t keep) {
    const uint32_t c = *s_cnt;
    const uint32_t n = next_pow2(c < 2u ? 2u : c);
    for (uint32_t i = c + threadIdx.x; i < n; i += blockDim.x) buf[i] = kEmptyKey;
    __syncthreads();
    block_bitonic_sort(buf, n);
    const uint32_t kept = c < keep ? c : keep;
    uint64_t thr = kept == keep ? buf[keep - 1] : kEmptyKey;
    __syncthreads();
    if (threadIdx.x == 0) *s_cnt = kept;
    __syncthreads();
    return thr;
}

// Persistent: each workgroup repeatedly takes the next query from `order`
// (heaviest first) through a global atomic counter, so queries of very uneven
// cost (Zipf lists, per-query nprobe) are balanced dynamically.
template <bool LISTORDER, bool GBUF, int METRIC>
__global__ __launch_bounds__(kIvfThreads) void ivf_scan_kernel(
    IvfPqGpuTables t, const uint8_t* __restrict__ codes_lo, const float* __restrict__ Qm,
    uint32_t nq, const uint32_t* __restrict__ order, uint32_t* __restrict__ next_q,
    const uint32_t* __restrict__ probes, uint32_t max_nprobe, const uint32_t* __restrict__ k_q,
    uint32_t k_u, uint32_t shortlist_u, uint32_t cap, int do_rerank,
    const uint32_t* __restrict__ mask, uint64_t* __restrict__ gbuf,
    GpuSearchResult* __restrict__ out, uint32_t out_stride) {
    extern __shared__ uint8_t s_dyn[];
    float* lut = reinterpret_cast<float*>(s_dyn);
    uint64_t* buf = GBUF ? gbuf + (size_t)blockIdx.x * cap
                         : reinterpret_cast<uint64_t*>(s_dyn + (size_t)t.m_pq * 256 * sizeof(float));
    __shared__ uint32_t s_cnt;
    __shared__ uint64_t s_thr;
    __shared__ uint32_t s_item;

  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(next_q, 1u);
    __syncthreads();
    if (s_item >= nq) return;
    const uint32_t q = order[s_item];
    const uint32_t M = t.m_pq, sub = t.sub_dim, dim = t.dim;
    const float* qv = Qm + (size_t)q * dim;
    uint32_t k = k_q ? k_q[q] : k_u;
    if (k > out_stride) k = out_stride;
    uint32_t S = shortlist_u < k ? k : shortlist_u;
    if (S > cap - kEntriesPerCheck) S = cap - kEntriesPerCheck;
    if (threadIdx.x == 0) {
        s_cnt = 0;
        s_thr = kEmptyKey;
    }
    __syncthreads();

    for (uint32_t p = 0; p < max_nprobe; ++p) {
        const uint32_t c = probes[(size_t)q * max_nprobe + p];
        if (c >= t.nlist) break;  // probes are packed; the rest are empty
        const float* cen = t.d_centroids + (size_t)c * dim;
        __syncthreads();  // previous list's lookups are done with the LUT
        for (uint32_t idx = threadIdx.x; idx < M * 256u; idx += blockDim.x) {
            const uint32_t m = idx >> 8, j = idx & 255u;
            const float* cb = t.d_pq_codebooks + ((size_t)m * 256u + j) * sub;
            float dist = 0.0f;
            for (uint32_t s = 0; s < sub; ++s) {
                float r = qv[m * sub + s] - cen[m * sub + s];
                float diff = r - cb[s];
                dist = __builtin_fmaf(diff, diff, dist);
            }
            lut[idx] = dist;
        }
        __syncthreads();
        const uint32_t beg = t.d_ivf_offsets[c], end = t.d_ivf_offsets[c + 1];
        for (uint32_t base = beg; base < end; base += kEntriesPerCheck) {
            const uint64_t thr = s_thr;
            if (M == 16u) {
                // Issue the id and code loads of all passes first so that each
                // wave keeps kPassesPerCheck 128-bit loads in flight.
                uint32_t ids[kPassesPerCheck];
                uint4 cw[kPassesPerCheck];
#pragma unroll
                for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass) {
                    const uint32_t e = base + pass * kIvfThreads + threadIdx.x;
                    const uint32_t ec = e < end ? e : end - 1;
                    ids[pass] = t.d_ivf_vec_ids[ec];
                    if (LISTORDER) cw[pass] = *reinterpret_cast<const uint4*>(codes_lo + (size_t)ec * 16u);
                }
                if (!LISTORDER) {
#pragma unroll
                    for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass) {
                        const uint8_t* cp = t.d_pq_codes + (size_t)ids[pass] * 16u;
                        if ((reinterpret_cast<uintptr_t>(cp) & 15u) == 0u) {
                            cw[pass] = *reinterpret_cast<const uint4*>(cp);
                        } else {
                            uint32_t w4[4];
                            for (int b = 0; b < 4; ++b)
                                w4[b] = cp[4 * b] | (cp[4 * b + 1] << 8) | (cp[4 * b + 2] << 16) | ((uint32_t)cp[4 * b + 3] << 24);
                            cw[pass] = make_uint4(w4[0], w4[1], w4[2], w4[3]);
                        }
                    }
                }
#pragma unroll
                for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass) {
                    const uint32_t e = base + pass * kIvfThreads + threadIdx.x;
                    const uint32_t id = ids[pass];
                    if (e >= end || !mask_allows(mask, id)) continue;
                    const uint32_t ww[4] = {cw[pass].x, cw[pass].y, cw[pass].z, cw[pass].w};
                    float adc = 0.0f;
#pragma unroll
                    for (int b = 0; b < 16; ++b) {
                        const uint32_t code = (ww[b >> 2] >> (8 * (b & 3))) & 255u;
                        adc = adc + lut[b * 256u + code];
                    }
                    if (adc < kAdmit) {
                        const uint64_t key = make_key(adc, id);
                        if (key < thr) buf[atomicAdd(&s_cnt, 1u)] = key;
                    }
                }
            } else {
#pragma unroll 1
                for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass) {
                    const uint32_t e = base + pass * kIvfThreads + threadIdx.x;
                    if (e >= end) break;
                    const uint32_t id = t.d_ivf_vec_ids[e];
                    if (!mask_allows(mask, id)) continue;
                    const uint8_t* cp = LISTORDER ? codes_lo + (size_t)e * M : t.d_pq_codes + (size_t)id * M;
                    float adc = 0.0f;
                    for (uint32_t m = 0; m < M; ++m) adc = adc + lut[m * 256u + cp[m]];
                    if (adc < kAdmit) {
                        const uint64_t key = make_key(adc, id);
                        if (key < thr) buf[atomicAdd(&s_cnt, 1u)] = key;
                    }
                }
            }
            __syncthreads();
            if (s_cnt + kEntriesPerCheck > cap) {
                const uint64_t nt = block_prune(buf, &s_cnt, S);
                if (threadIdx.x == 0 && nt < s_thr) s_thr = nt;
                __syncthreads();
            }
        }
    }
    __syncthreads();
    block_prune(buf, &s_cnt, S);
    const uint32_t kept = s_cnt;
    // Exact rerank of the shortlist, in place.
    if (do_rerank && t.d_vectors) {
        for (uint32_t i = threadIdx.x; i < kept; i += blockDim.x) {
            const uint32_t id = key_slot(buf[i]);
            float d = id < t.num_vectors
                ? exact_distance<METRIC>(qv, t.d_vectors + (size_t)id * dim, dim)
                : FV_EMPTY_DIST;
            buf[i] = d < kAdmit ? make_key(d, id) : kEmptyKey;
        }
        const uint32_t n = next_pow2(kept < 2u ? 2u : kept);
        for (uint32_t i = kept + threadIdx.x; i < n; i += blockDim.x) buf[i] = kEmptyKey;
        __syncthreads();
        block_bitonic_sort(buf, n);
    }
    for (uint32_t i = threadIdx.x; i < out_stride; i += blockDim.x) {
        GpuSearchResult r;
        const uint64_t key = (i < k && i < kept) ? buf[i] : kEmptyKey;
        if (key == kEmptyKey) {
            r.id = FV_EMPTY_ID;
            r.distance = FV_EMPTY_DIST;
        } else {
            r.id = key_slot(key);
            r.distance = key_dist(key);
        }
        out[(size_t)q * out_stride + i] = r;
    }
  }
}

// Scan cost of each query (entries over its probed lists), bucketed by
// log2 so that the heaviest queries are dispatched first.
__global__ void ivf_cost_kernel(const uint32_t* __restrict__ probes, uint32_t max_nprobe,
                                const uint32_t* __restrict__ offsets, uint32_t nlist,
                                uint32_t nq, uint32_t* __restrict__ bucket,
                                uint32_t* __restrict__ bucket_count) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint64_t cost = 0;
    for (uint32_t p = 0; p < max_nprobe; ++p) {
        const uint32_t c = probes[(size_t)q * max_nprobe + p];
        if (c >= nlist) break;
        cost += offsets[c + 1] - offsets[c];
    }
    const uint32_t c32 = cost > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)cost;
    const uint32_t b = (uint32_t)__clz(c32 | 1u);  // 0 = heaviest
    bucket[q] = b;
    atomicAdd(&bucket_count[b], 1u);
}

__global__ void ivf_order_kernel(const uint32_t* __restrict__ bucket,
                                 const uint32_t* __restrict__ bucket_count, uint32_t nq,
                                 uint32_t* __restrict__ bucket_fill, uint32_t* __restrict__ order) {
    __shared__ uint32_t s_off[33];
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int b = 0; b < 33; ++b) {
            s_off[b] = run;
            run += bucket_count[b];
        }
    }
    __syncthreads();
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
        const uint32_t b = bucket[q];
        order[s_off[b] + atomicAdd(&bucket_fill[b], 1u)] = q;
    }
}

// Probe list of each query from the coarse top-k results: the first np_q
// cluster ids, empty past np_q.
__global__ void probes_from_results_kernel(const GpuSearchResult* __restrict__ res, uint32_t nq,
                                           uint32_t max_nprobe, const uint32_t* __restrict__ nprobe_q,
                                           uint32_t nprobe_u, uint32_t nlist,
                                           uint32_t* __restrict__ probes) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nq * max_nprobe) return;
    const uint32_t q = (uint32_t)(i / max_nprobe), p = (uint32_t)(i % max_nprobe);
    uint32_t np = nprobe_q ? nprobe_q[q] : nprobe_u;
    if (np == 0) np = 1;
    if (np > nlist) np = nlist;
    probes[i] = p < np ? res[i].id : FV_EMPTY_ID;
}

__global__ void list_order_codes_kernel(const uint8_t* __restrict__ codes,
                                        const uint32_t* __restrict__ vec_ids, uint32_t n,
                                        uint32_t M, uint8_t* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * M) return;
    size_t e = i / M, m = i % M;
    out[i] = codes[(size_t)vec_ids[e] * M + m];
}

IvfPqPlan plan_ivfpq(const IvfPqGpuTables* t, uint32_t nq, uint32_t max_nprobe,
                     uint32_t shortlist, uint32_t max_k) {
    IvfPqPlan p{};
    uint32_t S = shortlist < max_k ? max_k : shortlist;
    p.shortlist = S;
    p.max_nprobe = max_nprobe;
    p.cap = next_pow2(S + kEntriesPerCheck);
    p.npow = next_pow2(t->nlist < 2u ? 2u : t->nlist);
    const size_t lut_bytes = (size_t)t->m_pq * 256 * sizeof(float);
    p.gbuf = lut_bytes + (size_t)p.cap * 8 > 64 * 1024;
    p.lds_scan = lut_bytes + (p.gbuf ? 0 : (size_t)p.cap * 8);
    p.grid = nq < kMaxScanGrid ? nq : kMaxScanGrid;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    p.off_probes = take((size_t)nq * max_nprobe * 4);
    p.off_bucket = take((size_t)nq * 4);
    p.off_order = take((size_t)nq * 4);
    p.off_counters = take(68 * 4);  // 33 counts, 33 fills, next query, pad
    p.off_gbuf = take(p.gbuf ? (size_t)p.grid * p.cap * 8 : 0);
    p.tiled_coarse = max_nprobe <= 256u && t->nlist >= 64u;
    if (p.tiled_coarse) {
        p.coarse_plan = plan_flat_exact(t->nlist, t->dim, nq, max_nprobe, METRIC_L2);
        p.off_coarse_res = take((size_t)nq * max_nprobe * sizeof(GpuSearchResult));
        p.off_coarse_ws = take(p.coarse_plan.workspace_bytes);
    }
    p.workspace_bytes = off;
    return p;
}

template <bool LO, bool GB>
static void launch_scan_m(int metric, dim3 g, size_t lds, hipStream_t s, const IvfPqGpuTables& t,
                          const uint8_t* lo, const float* Qm, uint32_t nq, const uint32_t* order,
                          uint32_t* next_q, const uint32_t* probes, uint32_t mnp,
                          const uint32_t* k_q, uint32_t k_u, uint32_t S, uint32_t cap, int rr,
                          const uint32_t* mask, uint64_t* gbuf, GpuSearchResult* out,
                          uint32_t stride) {
    if (metric == METRIC_INNER_PRODUCT)
        ivf_scan_kernel<LO, GB, METRIC_INNER_PRODUCT><<<g, kIvfThreads, lds, s>>>(t, lo, Qm, nq, order, next_q, probes, mnp, k_q, k_u, S, cap, rr, mask, gbuf, out, stride);
    else if (metric == METRIC_COSINE)
        ivf_scan_kernel<LO, GB, METRIC_COSINE><<<g, kIvfThreads, lds, s>>>(t, lo, Qm, nq, order, next_q, probes, mnp, k_q, k_u, S, cap, rr, mask, gbuf, out, stride);
    else
        ivf_scan_kernel<LO, GB, METRIC_L2><<<g, kIvfThreads, lds, s>>>(t, lo, Qm, nq, order, next_q, probes, mnp, k_q, k_u, S, cap, rr, mask, gbuf, out, stride);
}

hipError_t launch_ivfpq(const IvfPqGpuTables* t, const uint8_t* codes_lo, const float* Qm,
                        uint32_t nq, const uint32_t* k_q, uint32_t k_u, const uint32_t* nprobe_q,
                        uint32_t nprobe_u, int do_rerank, int metric, const uint32_t* mask,
                        GpuSearchResult* out, uint32_t out_stride, void* workspace,
                        const IvfPqPlan& p, hipStream_t stream) {
    char* ws = static_cast<char*>(workspace);
    uint32_t* probes = reinterpret_cast<uint32_t*>(ws + p.off_probes);
    uint64_t* gbuf = p.gbuf ? reinterpret_cast<uint64_t*>(ws + p.off_gbuf) : nullptr;
    uint32_t* bucket = reinterpret_cast<uint32_t*>(ws + p.off_bucket);
    uint32_t* order = reinterpret_cast<uint32_t*>(ws + p.off_order);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(ws + p.off_counters);
    if (p.tiled_coarse) {
        // Tiled exact scan with fmaf accumulation: same values and (distance,
        // cluster) order as the per-query coarse kernel, at GEMM-like reuse.
        GpuSearchResult* cres = reinterpret_cast<GpuSearchResult*>(ws + p.off_coarse_res);
        const hipError_t ce = launch_flat_exact(t->d_centroids, nullptr, t->nlist, t->dim, Qm, nq,
                                                p.max_nprobe, METRIC_L2, nullptr, cres, p.max_nprobe,
                                                ws + p.off_coarse_ws, p.coarse_plan, stream, true);
        if (ce != hipSuccess) return ce;
        const size_t tot = (size_t)nq * p.max_nprobe;
        probes_from_results_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, stream>>>(
            cres, nq, p.max_nprobe, nprobe_q, nprobe_u, t->nlist, probes);
    } else {
        ivf_coarse_kernel<<<nq, 256, (size_t)p.npow * 8, stream>>>(
            t->d_centroids, t->nlist, t->dim, Qm, nprobe_q, nprobe_u, p.max_nprobe, p.npow, probes);
    }
    fill_u32(ctr, 68, 0u, stream);
    ivf_cost_kernel<<<(nq + 255) / 256, 256, 0, stream>>>(probes, p.max_nprobe, t->d_ivf_offsets,
                                                         t->nlist, nq, bucket, ctr);
    ivf_order_kernel<<<(nq + 255) / 256 < 64 ? (nq + 255) / 256 : 64, 256, 0, stream>>>(
        bucket, ctr, nq, ctr + 33, order);
    uint32_t* next_q = ctr + 66;
    void* timer = timer_begin(stream);
    const dim3 g(p.grid);
    const bool lo = codes_lo != nullptr;
    if (lo && p.gbuf)
        launch_scan_m<true, true>(metric, g, p.lds_scan, stream, *t, codes_lo, Qm, nq, order, next_q, probes, p.max_nprobe, k_q, k_u, p.shortlist, p.cap, do_rerank, mask, gbuf, out, out_stride);
    else if (lo)
        launch_scan_m<true, false>(metric, g, p.lds_scan, stream, *t, codes_lo, Qm, nq, order, next_q, probes, p.max_nprobe, k_q, k_u, p.shortlist, p.cap, do_rerank, mask, gbuf, out, out_stride);
    else if (p.gbuf)
        launch_scan_m<false, true>(metric, g, p.lds_scan, stream, *t, codes_lo, Qm, nq, order, next_q, probes, p.max_nprobe, k_q, k_u, p.shortlist, p.cap, do_rerank, mask, gbuf, out, out_stride);
    else
        launch_scan_m<false, false>(metric, g, p.lds_scan, stream, *t, codes_lo, Qm, nq, order, next_q, probes, p.max_nprobe, k_q, k_u, p.shortlist, p.cap, do_rerank, mask, gbuf, out, out_stride);
    timer_end(timer, stream);
    return hipGetLastError();
}

static bool tables_ok(const IvfPqGpuTables* t) {
    return t->d_centroids && t->d_pq_codebooks && t->d_pq_codes && t->d_ivf_offsets &&
           t->d_ivf_vec_ids && t->dim > 0 && t->m_pq > 0 && t->sub_dim > 0 &&
           t->m_pq * t->sub_dim == t->dim && t->nlist <= kMaxNlist;
}

}  // namespace fv

extern "C" {

// cuda_bridge.h:96-107 / kernels/src/ivf_pq_lookup.cu:339-473.
int cuda_ivf_pq_search_batch(const IvfPqGpuTables* tables, const float* d_queries,
                             uint32_t num_queries, uint32_t top_k, uint32_t nprobe,
                             uint32_t rerank_k, MetricType metric, const uint32_t* d_valid_mask,
                             GpuSearchResult* d_out_results, void* stream) {
    if (!tables || !d_queries || !d_out_results) return -1;
    if (num_queries == 0 || top_k == 0 || tables->m_pq == 0 || tables->nlist == 0) return 0;
    if (!fv::tables_ok(tables)) return (int)hipErrorInvalidValue;
    if (top_k > 256u) top_k = 256u;
    uint32_t np = nprobe == 0 ? 1u : nprobe;
    if (np > tables->nlist) np = tables->nlist;
    uint32_t S = rerank_k == 0 ? top_k : rerank_k;
    if (S < top_k) S = top_k;
    if (S > 10240u) S = 10240u;
    const int do_rerank = rerank_k != 1u && tables->d_vectors != nullptr;
    hipStream_t s = static_cast<hipStream_t>(stream);
    fv::IvfPqPlan p = fv::plan_ivfpq(tables, num_queries, np, S, top_k);
    void* ws = nullptr;
    hipError_t e = hipMallocAsync(&ws, p.workspace_bytes, s);
    if (e != hipSuccess) return (int)e;
    e = fv::launch_ivfpq(tables, nullptr, d_queries, num_queries, nullptr, top_k, nullptr, np,
                         do_rerank, (int)metric, d_valid_mask, d_out_results, top_k, ws, p, s);
    hipError_t e2 = hipFreeAsync(ws, s);
    if (e != hipSuccess) return (int)e;
    return e2 == hipSuccess ? 0 : (int)e2;
}

int fv_ivf_pq_list_order_codes(const IvfPqGpuTables* tables, uint8_t* d_codes_out, void* stream) {
    if (!tables || !d_codes_out || !tables->d_pq_codes || !tables->d_ivf_vec_ids) return -1;
    size_t total = (size_t)tables->num_vectors * tables->m_pq;
    if (total == 0) return 0;
    fv::list_order_codes_kernel<<<(unsigned)((total + 255) / 256), 256, 0,
                                  static_cast<hipStream_t>(stream)>>>(
        tables->d_pq_codes, tables->d_ivf_vec_ids, tables->num_vectors, tables->m_pq, d_codes_out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

size_t fv_ivf_pq_workspace_size(const IvfPqGpuTables* tables, uint32_t num_queries,
                                uint32_t max_nprobe, uint32_t shortlist, uint32_t max_top_k) {
    if (!tables || max_top_k == 0 || max_top_k > FV_MAX_TOPK || shortlist > FV_MAX_SHORTLIST ||
        max_nprobe == 0 || tables->nlist == 0)
        return 0;
    uint32_t np = max_nprobe > tables->nlist ? tables->nlist : max_nprobe;
    return fv::plan_ivfpq(tables, num_queries, np, shortlist, max_top_k).workspace_bytes;
}

int fv_ivf_pq_search(const IvfPqGpuTables* tables, const uint8_t* d_codes_list_order,
                     const float* d_queries, uint32_t num_queries, const uint32_t* d_top_k,
                     uint32_t max_top_k, const uint32_t* d_nprobe, uint32_t max_nprobe,
                     uint32_t shortlist, int rerank, MetricType metric,
                     const uint32_t* d_valid_mask, GpuSearchResult* d_out_results,
                     void* d_workspace, size_t workspace_bytes, void* stream) {
    if (!tables || !d_queries || !d_out_results) return -1;
    if (!fv::tables_ok(tables) || max_top_k == 0 || max_top_k > FV_MAX_TOPK ||
        shortlist > FV_MAX_SHORTLIST || max_nprobe == 0)
        return (int)hipErrorInvalidValue;
    if (num_queries == 0) return 0;
    uint32_t np = max_nprobe > tables->nlist ? tables->nlist : max_nprobe;
    fv::IvfPqPlan p = fv::plan_ivfpq(tables, num_queries, np, shortlist, max_top_k);
    if (!d_workspace || workspace_bytes < p.workspace_bytes) return (int)hipErrorInvalidValue;
    const int do_rerank = rerank && tables->d_vectors != nullptr;
    hipError_t e = fv::launch_ivfpq(tables, d_codes_list_order, d_queries, num_queries, d_top_k,
                                    max_top_k, d_nprobe, np, do_rerank, (int)metric, d_valid_mask,
                                    d_out_results, max_top_k, d_workspace, p,
                                    static_cast<hipStream_t>(stream));
    return e == hipSuccess ? 0 : (int)e;
}

}  // extern "C"
