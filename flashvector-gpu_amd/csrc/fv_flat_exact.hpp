// Internal launch plan of the exact FP32 flat scan.
#pragma once

#include "../../include/flashvector/cuda_bridge.h"
#include "../../include/flashvector/fv_ext.h"
#include "fv_device.hpp"

namespace fv {

struct FlatExactPlan {
    uint32_t keep;            // results kept per (chunk, query)
    uint32_t cap;             // candidate buffer capacity per (chunk, query)
    uint32_t rows_per_chunk;  // multiple of the 64-row tile
    uint32_t n_chunks;
    uint32_t qblocks;
    size_t off_cand, off_partial, off_norms;
    size_t workspace_bytes;
};

// max_bytes: fewer row chunks when the plan's workspace would exceed it
// (workspace_bytes = 0 when even one chunk does not fit).
FlatExactPlan plan_flat_exact(uint32_t n, uint32_t dim, uint32_t nq, uint32_t top_k, int metric,
                              size_t max_bytes = ~(size_t)0);

hipError_t launch_flat_exact(const float* X, const uint32_t* vec_ids, uint32_t n, uint32_t dim,
                             const float* Qm, uint32_t nq, uint32_t top_k, int metric,
                             const uint32_t* mask, GpuSearchResult* out, uint32_t out_stride,
                             void* workspace, const FlatExactPlan& p, hipStream_t stream,
                             bool fma_l2 = false);

// Kernel timer hooks (fv_kernel_timer_enable): begin returns a token or NULL.
void* timer_begin(hipStream_t stream);
void timer_end(void* token, hipStream_t stream);

// Device fill of n words (a kernel, so calls stay graph-capturable).
void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t stream);

hipError_t fill_results(GpuSearchResult* out, size_t count, uint32_t id_bits, uint32_t dist_bits,
                        hipStream_t stream);

// Top-k merge of num_shards lists per query (fv_merge.hip), out_k <= 1024.
hipError_t launch_merge_topk(const GpuSearchResult* in, uint32_t num_shards, uint32_t nq,
                             uint32_t shard_k, uint32_t out_k, GpuSearchResult* out,
                             hipStream_t stream);

}  // namespace fv
