/*
 * C ABI of the query-scan hot path, MI355X (gfx950, HIP) implementation.
 *
 * The symbol names and signatures are the reference's, because its Rust FFI
 * binds them by name (crates/engine/src/ffi.rs:234-455); "cuda_" is part of
 * that ABI name, not a CUDA dependency. Every `void* stream` is a hipStream_t
 * (NULL = the default stream), every d_* pointer is device memory (or mapped
 * pinned host memory where noted), results are written asynchronously on the
 * stream. Return codes: 0 success, -1 NULL required pointer, otherwise a
 * hipError_t value. Out-of-range sizes are clamped as the reference does.
 */
#ifndef FLASHVECTOR_AMD_CUDA_BRIDGE_H
#define FLASHVECTOR_AMD_CUDA_BRIDGE_H

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device, memory, streams (replaces kernels/include/cuda_bridge.h:11-39,
 *      implemented in kernels/src/hnsw_traverse.cu:525-637) ---------------- */
int cuda_init_device(int device_id);                                   /* cuda_bridge.h:12 */
int cuda_set_device(int device_id);                                    /* cuda_bridge.h:13 */
int cuda_get_device_count(int* count);                                 /* cuda_bridge.h:14 */
int cuda_get_device_memory(size_t* free_bytes, size_t* total_bytes);   /* cuda_bridge.h:15 */
int cuda_device_synchronize(void);                                     /* cuda_bridge.h:16 */
int cuda_enable_peer_access(int device, int peer);                     /* cuda_bridge.h:17 */
int cuda_memcpy_peer_async(void* dst, int dst_device, const void* src, /* cuda_bridge.h:18-25 */
                           int src_device, size_t bytes, void* stream);
int cuda_malloc_device(void** ptr, size_t bytes);                      /* cuda_bridge.h:28 */
int cuda_free_device(void* ptr);                                       /* cuda_bridge.h:29 */
int cuda_malloc_host(void** ptr, size_t bytes);  /* mapped pinned;       cuda_bridge.h:30 */
int cuda_free_host(void* ptr);                                         /* cuda_bridge.h:31 */
int cuda_memcpy_h2d_async(void* dst, const void* src, size_t bytes, void* stream); /* :32 */
int cuda_memcpy_d2h_async(void* dst, const void* src, size_t bytes, void* stream); /* :33 */
int cuda_memset_device_async(void* dst, int value, size_t bytes, void* stream);    /* :34 */
int cuda_host_get_device_pointer(void** dev_ptr, void* host_ptr);                  /* :35 */
int cuda_create_stream(void** stream);                                 /* cuda_bridge.h:38 */
int cuda_destroy_stream(void* stream);                                 /* cuda_bridge.h:39 */
int cuda_sync_stream(void* stream);                                    /* cuda_bridge.h:40 */

/* ---- IVF-PQ ADC scan + shortlist + exact FP32 rerank.
 * Replaces cuda_bridge.h:96-107 (kernels/src/ivf_pq_lookup.cu:339-473).
 * top_k <= 256; nprobe clamped to [1, nlist]; shortlist S = rerank_k ? rerank_k
 * : top_k, raised to top_k, capped at 10240; rerank unless rerank_k == 1 or
 * tables->d_vectors is NULL. Ties go to the lower vector id. */
int cuda_ivf_pq_search_batch(const IvfPqGpuTables* tables, const float* d_queries,
                             uint32_t num_queries, uint32_t top_k, uint32_t nprobe,
                             uint32_t rerank_k, MetricType metric, const uint32_t* d_valid_mask,
                             GpuSearchResult* d_out_results, void* stream);

/* ---- FP16 matrix-core flat scan. Replaces cuda_bridge.h:138-147
 * (kernels/src/ivf_quant_lookup.cu:887-1058,1106-1123). top_k <= 256. When
 * tables->d_vectors is non-NULL, a shortlist of 8*top_k is reranked exactly
 * in FP32 (the reference never reads d_vectors and returns FP16 scores).
 * Rows of up to 1536 padded dims are scored on the matrix cores; wider rows
 * need tables->d_vectors and are scanned exactly in FP32 (same k best).
 * The codes must be the binary16 encoding of d_vectors rounded to nearest
 * (crates/engine/src/quant.rs encode_fp16_padded): the exact rerank then
 * skips rows that this rounding bound proves cannot be among the k best. */
int cuda_fp16_tc_search_batch(const Fp16GpuTables* tables, const float* d_queries,
                              uint32_t num_queries, uint32_t top_k, MetricType metric,
                              const uint32_t* d_valid_mask, GpuSearchResult* d_out_results,
                              void* stream);

/* ---- Exact FP32 flat scan. Replaces cuda_bridge.h:150-162
 * (kernels/src/sparse_hybrid.cu:13-107,504-524). top_k clamped to 256 (also the
 * output row stride); mask tested on the slot; ids remapped through
 * d_vec_ids when non-NULL; num_vectors == 0 fills the output with 0xFF bytes. */
int cuda_flat_scan_batch(const float* d_vectors, const uint32_t* d_vec_ids, uint32_t num_vectors,
                         uint32_t dim, const float* d_queries, uint32_t num_queries,
                         uint32_t top_k, MetricType metric, const uint32_t* d_valid_mask,
                         GpuSearchResult* d_out_results, void* stream);

/* ---- Top-k merge across shards (or main + delta). Replaces
 * cuda_bridge.h:165-173 (kernels/src/sparse_hybrid.cu:139-218,527-542).
 * Input [num_shards][num_queries][shard_top_k]; the first 4096 entries per
 * query are considered; ordered by (distance, id), first occurrence of an id
 * kept; out_top_k clamped to 256. */
int cuda_merge_topk(const GpuSearchResult* d_shard_results, uint32_t num_shards,
                    uint32_t num_queries, uint32_t shard_top_k, uint32_t out_top_k,
                    GpuSearchResult* d_out_results, void* stream);

/* ---- Known-answer-test kernels. Replace cuda_bridge.h:203-221
 * (kernels/src/hnsw_traverse.cu:387-437,694-733). The bitonic test sorts up
 * to 64 pairs (one wavefront). */
int cuda_compute_distances_warp(const float* d_queries, const float* d_dataset,
                                uint32_t num_queries, uint32_t num_vectors, uint32_t dim,
                                MetricType metric, float* d_out_distances, void* stream);
int cuda_bitonic_sort_test(const float* d_keys_in, const uint32_t* d_vals_in, float* d_keys_out,
                           uint32_t* d_vals_out, uint32_t n, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FLASHVECTOR_AMD_CUDA_BRIDGE_H */
