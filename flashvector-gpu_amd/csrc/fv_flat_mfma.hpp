// Internal launch plan of the reduced-precision (matrix-core) flat scan.
#pragma once

#include "fv_flat_exact.hpp"

namespace fv {

// Per-call arguments shared by the scan and the select/rerank kernels.
struct FlatSearchArgs {
    const uint32_t* mask;     // validity words or NULL
    const uint32_t* d_top_k;  // per-query k or NULL
    uint32_t top_k;           // uniform k and output row stride
    uint32_t factor;          // shortlist = factor * k (0: fixed shortlist)
    uint32_t shortlist;       // fixed shortlist (0: 8 * k)
    uint32_t s_max;           // deepest shortlist of the call
    uint32_t dim;
    int metric;
};

struct FlatMfmaPlan {
    uint32_t s_max;           // deepest per-query shortlist
    uint32_t cap;             // candidate buffer capacity per (chunk, query)
    uint32_t prune_regs;      // keys per lane of a wave prune (4 or 16), 0 = workgroup prunes
    bool block_prune;         // prune with the whole workgroup (deep shortlists)
    size_t lds_scan;          // dynamic LDS of the scan kernel
    uint32_t lcap;            // LDS keys of the select/rerank kernel
    uint32_t ks;              // 32-wide K steps per row
    uint32_t qblocks;         // 256-query blocks
    uint32_t n_chunks;
    bool compact;             // scan the compacted list of valid rows
    uint32_t bound_chunks;    // first-pass chunks (0 = no first pass)
    uint32_t bound_rows;      // rows covered by the first pass
    uint32_t mask_blocks;
    size_t off_gthr, off_counts, off_cand, off_qh, off_rows, off_nrows, off_blocks, off_minima;
    size_t workspace_bytes;
};

FlatMfmaPlan plan_flat_mfma(uint32_t n, uint32_t padded_dim, uint32_t nq, const FlatSearchArgs& a,
                            bool has_mask, uint32_t valid_count);

hipError_t launch_flat_mfma(const FvFlatTables* t, const float* Qm, uint32_t nq,
                            const FlatSearchArgs& a, GpuSearchResult* out, void* workspace,
                            const FlatMfmaPlan& p, hipStream_t stream);

}  // namespace fv
