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
    uint32_t flags;           // FV_SEARCH_* bits
    uint32_t n_chunks;        // plan override (0 = automatic)
    uint32_t bound_div;       // plan override (0 = automatic, ~0u = no first pass)
    // Row count known only on the device (mask compacted without a host
    // count): the main pass launches the plan's chunk count as an upper bound
    // and uses min(chunk_want, ceil(rows / chunk_min_rows)) chunks (a
    // multiple of 8 from 8 on); surplus workgroups only zero their counts.
    // 0 = the launched chunk count is used as is.
    uint32_t chunk_want;
    uint32_t chunk_min_rows;
    // Caller's sum of the per-query k (host hint for sizing per-query
    // candidate buffers; 0 = unknown).
    uint32_t top_k_sum;
    // Per-query candidate buffers (workgroup-prune scans only): in chunk c,
    // query q's buffer starts at key c * cand_stride + cand_off[q] and holds
    // cand_off[q + 1] - cand_off[q] keys. NULL: every buffer holds `cap` keys
    // at (c * nq + q) * cap.
    const uint32_t* cand_off;
    uint64_t cand_stride;
};

// Chunks of the main pass for `rows` rows (the planner's rule).
__host__ __device__ inline uint32_t chunks_for_rows(uint32_t rows, uint32_t want,
                                                    uint32_t min_rows, uint32_t cap) {
    const uint32_t by_rows = (rows + min_rows - 1) / min_rows;
    uint32_t c = want < by_rows ? want : by_rows;
    if (c > 256u) c = 256u;
    if (c >= 8u) c &= ~7u;
    if (c < 1u) c = 1u;
    return c < cap ? c : cap;
}

struct FlatMfmaPlan {
    uint32_t s_max;           // deepest per-query shortlist
    uint32_t cap;             // candidate buffer capacity per (chunk, query)
    uint32_t prune_regs;      // keys per lane of a wave prune (16), 0 = workgroup prunes
    bool block_prune;         // prune with the whole workgroup (deep shortlists)
    size_t lds_scan;          // dynamic LDS of the scan kernel
    uint32_t lcap;            // LDS keys of the select/rerank kernel
    uint32_t sel_batch;       // threads of the select kernel (candidates per round)
    uint32_t final_threads;   // threads of the finalize kernel
    size_t lds_select;        // dynamic LDS of the select kernel
    bool tail;                // fused wave-per-query select/rerank/finalize
    uint32_t tail_lw;         // its LDS keys per wavefront (next_pow2(s_max + 64))
    uint32_t tail_hw;         // its rerank slots per wavefront (next_pow2(s_max))
    uint32_t tail_wbytes;     // its LDS bytes per wavefront
    size_t lds_tail;          // its dynamic LDS per workgroup
    uint32_t hcap;            // keys of the finalize kernel (next_pow2(s_max))
    size_t lds_final;         // dynamic LDS of the finalize kernel
    uint32_t ks;              // K steps per row (32 elements for 16-bit, 128 for FP8)
    uint32_t qw;              // queries per wave
    uint32_t qb;              // queries per workgroup (4 * qw)
    uint32_t qrow_bytes;      // bytes of one encoded query row
    uint32_t qblocks;         // query blocks
    uint32_t n_chunks;
    bool compact;             // scan the compacted list of valid rows
    bool dyn_rows;            // compacted without a host count: chunks chosen on the device
    uint32_t chunk_want;      // device-chosen chunks: FlatSearchArgs::chunk_want / chunk_min_rows
    uint32_t chunk_min_rows;
    uint32_t bound_chunks;    // first-pass chunks (0 = no first pass)
    uint32_t bound_rows;      // rows covered by the first pass
    uint32_t mask_blocks;
    bool percap;              // per-query candidate buffers (FlatSearchArgs::cand_off)
    uint32_t slack;           // their keys beyond the query's shortlist
    uint64_t cand_keys;       // candidate keys per chunk
    size_t off_candoff;
    size_t off_gthr, off_counts, off_cand, off_qh, off_qscale, off_rows, off_nrows, off_blocks,
        off_minima, off_sel, off_seln, off_seloff, off_err;
    size_t workspace_bytes;
};

// Operands of one scan launch (main pass and optional first pass).
struct ScanLaunch {
    const FlatMfmaPlan* p;
    const uint8_t* X;          // row codes
    uint32_t row_bytes;        // stride of X in bytes
    const float* norms;
    const float* row_scales;   // FP8 only
    const uint8_t* Qh;         // encoded negated queries
    const float* qscale;       // FP8 only
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

FlatMfmaPlan plan_flat_mfma(uint32_t n, uint32_t padded_dim, int dtype, uint32_t nq,
                            const FlatSearchArgs& a, bool has_mask, uint32_t valid_count);

hipError_t launch_flat_mfma(const FvFlatTables* t, const float* Qm, uint32_t nq,
                            const FlatSearchArgs& a, GpuSearchResult* out, void* workspace,
                            const FlatMfmaPlan& p, hipStream_t stream);

// FP8 e4m3 rows with per-row scales (s = max|x|/448, code = e4m3(sign*x/s)),
// rows >= n written as zeros with scale 1 (fv_flat_mfma.hip).
void launch_encode_fp8_rows(const float* x, uint32_t n, uint32_t dim, uint32_t out_stride,
                            uint32_t total_rows, float sign, uint8_t* out, float* scales,
                            hipStream_t stream);

// Plan of fv_flat_search for these arguments (false if they are invalid).
bool flat_plan_for(const FvFlatTables* t, uint32_t nq, const FvFlatSearchParams* sp,
                   FlatMfmaPlan* out);

// First-pass bound selection (fv_flat_mfma.hip), called by the scan launchers.
void launch_bound_select(const ScanLaunch& L);

// Scan launchers, one per dtype and K-step range (fv_flat_mfma_*.hip).
hipError_t scan_f16_small(const ScanLaunch& L);   // ks in {1,2,3,4,6,8}
hipError_t scan_f16_large(const ScanLaunch& L);   // ks in {12,16,24}
hipError_t scan_f16_wide(const ScanLaunch& L);    // ks in {32,48}
hipError_t scan_bf16_small(const ScanLaunch& L);
hipError_t scan_bf16_large(const ScanLaunch& L);
hipError_t scan_bf16_wide(const ScanLaunch& L);
hipError_t scan_fp8(const ScanLaunch& L);         // ks in {1,2,4,6}
hipError_t scan_fp8_wide(const ScanLaunch& L);    // ks in {8,12,16,24}

}  // namespace fv
