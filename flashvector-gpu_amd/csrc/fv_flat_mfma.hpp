// Internal launch plan of the reduced-precision (matrix-core) flat scan.
#pragma once

#include "fv_flat_exact.hpp"

namespace fv {

struct FlatMfmaPlan {
    uint32_t keep;            // shortlist depth S kept per (chunk, query)
    uint32_t lpow;            // next power of two >= keep
    uint32_t cap;             // candidate buffer capacity per (chunk, query)
    uint32_t ks;              // 32-wide K steps per row
    uint32_t qblocks;         // 256-query blocks
    uint32_t rows_per_chunk;  // multiple of the 64-row tile
    uint32_t n_chunks;
    size_t off_gthr, off_cand, off_partial;
    size_t workspace_bytes;
};

FlatMfmaPlan plan_flat_mfma(uint32_t n, uint32_t padded_dim, uint32_t nq, uint32_t top_k,
                            uint32_t shortlist);

hipError_t launch_flat_mfma(const FvFlatTables* t, const float* Qm, uint32_t nq, uint32_t top_k,
                            int metric, const uint32_t* mask, GpuSearchResult* out,
                            uint32_t out_stride, void* workspace, const FlatMfmaPlan& p,
                            hipStream_t stream);

}  // namespace fv
