// Top-k merge across shards or main + delta lists (cuda_merge_topk).
//
// Reference: kernels/src/sparse_hybrid.cu:139-218 (distance-only bitonic sort,
// then a thread-0 serial de-duplication) and the host rule of
// crates/engine/src/index.rs:2637-2678 (first occurrence of an id in ascending
// distance order wins). Here de-duplication is a second sort: ordering by
// (id, distance) puts each id's best entry first, the survivors are re-sorted
// by (distance, id), and the first out_top_k are written.
#include "fv_flat_exact.hpp"

namespace fv {

constexpr uint32_t kMergeCap = 4096u;  // entries per query considered (sparse_hybrid.cu:211-212)

__global__ __launch_bounds__(256) void merge_topk_kernel(const GpuSearchResult* __restrict__ in,
                                                         uint32_t num_shards, uint32_t nq,
                                                         uint32_t shard_k, uint32_t total,
                                                         uint32_t n_pow2, uint32_t out_k,
                                                         GpuSearchResult* __restrict__ out) {
    extern __shared__ uint64_t s_m[];
    uint64_t* by_id = s_m;
    uint64_t* by_dist = s_m + n_pow2;
    const uint32_t q = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < n_pow2; i += blockDim.x) {
        uint64_t k = kEmptyKey;
        if (i < total) {
            uint32_t s = i / shard_k, j = i % shard_k;
            GpuSearchResult r = in[((size_t)s * nq + q) * shard_k + j];
            if (r.id != FV_EMPTY_ID) k = ((uint64_t)r.id << 32) | ord_f32(r.distance);
        }
        by_id[i] = k;
    }
    __syncthreads();
    block_bitonic_sort(by_id, n_pow2);
    for (uint32_t i = threadIdx.x; i < n_pow2; i += blockDim.x) {
        uint64_t k = by_id[i];
        bool first = k != kEmptyKey && (i == 0 || (by_id[i - 1] >> 32) != (k >> 32));
        by_dist[i] = first ? ((k << 32) | (k >> 32)) : kEmptyKey;
    }
    __syncthreads();
    block_bitonic_sort(by_dist, n_pow2);
    for (uint32_t i = threadIdx.x; i < out_k; i += blockDim.x) {
        uint64_t k = i < n_pow2 ? by_dist[i] : kEmptyKey;
        GpuSearchResult r;
        if (k == kEmptyKey) {
            r.id = FV_EMPTY_ID;
            r.distance = FV_EMPTY_DIST;
        } else {
            r.id = (uint32_t)k;
            r.distance = unord_f32((uint32_t)(k >> 32));
        }
        out[(size_t)q * out_k + i] = r;
    }
}

// Merge without the reference entry point's 256 clamp (out_k <= 1024 while
// num_shards * shard_k <= 4096); used by the host engine's main + delta merge.
hipError_t launch_merge_topk(const GpuSearchResult* in, uint32_t num_shards, uint32_t nq,
                             uint32_t shard_k, uint32_t out_k, GpuSearchResult* out,
                             hipStream_t stream) {
    if (num_shards == 0 || nq == 0 || out_k == 0) return hipSuccess;
    uint64_t total64 = (uint64_t)num_shards * shard_k;
    uint32_t total = total64 > kMergeCap ? kMergeCap : (uint32_t)total64;
    uint32_t n_pow2 = next_pow2(total < 2u ? 2u : total);
    merge_topk_kernel<<<nq, 256, (size_t)2 * n_pow2 * sizeof(uint64_t), stream>>>(
        in, num_shards, nq, shard_k ? shard_k : 1u, total, n_pow2, out_k, out);
    return hipGetLastError();
}

}  // namespace fv

extern "C" int cuda_merge_topk(const GpuSearchResult* d_shard_results, uint32_t num_shards,
                               uint32_t num_queries, uint32_t shard_top_k, uint32_t out_top_k,
                               GpuSearchResult* d_out_results, void* stream) {
    if (!d_shard_results || !d_out_results) return -1;
    if (num_shards == 0 || num_queries == 0 || out_top_k == 0) return 0;
    if (out_top_k > 256u) out_top_k = 256u;
    uint64_t total64 = (uint64_t)num_shards * shard_top_k;
    uint32_t total = total64 > fv::kMergeCap ? fv::kMergeCap : (uint32_t)total64;
    uint32_t n_pow2 = fv::next_pow2(total < 2u ? 2u : total);
    size_t lds = (size_t)2 * n_pow2 * sizeof(uint64_t);
    fv::merge_topk_kernel<<<num_queries, 256, lds, static_cast<hipStream_t>(stream)>>>(
        d_shard_results, num_shards, num_queries, shard_top_k ? shard_top_k : 1u, total, n_pow2,
        out_top_k, d_out_results);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}
