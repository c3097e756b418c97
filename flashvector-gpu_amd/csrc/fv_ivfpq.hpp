// Internal launch plan of the IVF-PQ scan.
#pragma once

#include "fv_flat_exact.hpp"

namespace fv {

// Coarse probe: the tiled exact scan for up to 1024 probes (at least 64
// lists), else every centroid's key sorted in LDS (up to 16384 lists, 128 KB),
// else the keys in the workspace and a radix selection there.
constexpr uint32_t kTiledCoarseMaxProbes = 1024;
constexpr uint32_t kLdsCoarseMaxLists = 16384;
enum CoarseMode : uint32_t { kCoarseTiled = 0, kCoarseLds = 1, kCoarseGlobal = 2 };
constexpr uint32_t kMaxScanGrid = 2048; // persistent scan workgroups (8 per CU on 256 CUs)
// Work items: a query of more than C probed entries is split into items of
// at most C, C = max(chunk_min, total / target_items, 2 * total / max_slots).
constexpr uint32_t kDefaultTargetItems = 1024;  // about the resident scan workgroups
constexpr uint32_t kDefaultChunkMin = 16384;
constexpr uint32_t kMaxSlots = 16384;             // partial shortlists of split queries
constexpr size_t kMaxPartBytes = 256u << 20;      // their keys, at most

struct IvfPqPlan {
    uint32_t shortlist;   // S, the ADC shortlist depth (>= every k)
    uint32_t max_nprobe;
    uint32_t cap;         // candidate buffer capacity per query
    uint32_t npow;        // next power of two >= nlist
    bool gbuf;            // candidate buffer in global memory (deep shortlists)
    size_t lds_scan;      // dynamic LDS of the scan kernel
    uint32_t grid;        // persistent scan workgroups
    uint32_t coarse_mode; // CoarseMode
    FlatExactPlan coarse_plan;
    size_t off_coarse_res, off_coarse_ws, off_coarse_keys;
    uint32_t max_slots;    // partial slots of split queries (0: no splitting)
    uint32_t target_items; // work items the batch is split into (0: one item per query)
    uint32_t chunk_min;    // smallest chunk of a split query, in entries
    size_t off_probes, off_bucket, off_order, off_counters, off_gbuf;
    size_t off_cost, off_items, off_slots, off_done, off_qbound, off_part_cnt, off_part_keys;
    size_t workspace_bytes;
};

IvfPqPlan plan_ivfpq(const IvfPqGpuTables* t, uint32_t nq, uint32_t max_nprobe,
                     uint32_t shortlist, uint32_t max_k);

hipError_t launch_ivfpq(const IvfPqGpuTables* t, const uint8_t* codes_lo, const float* Qm,
                        uint32_t nq, const uint32_t* k_q, uint32_t k_u, const uint32_t* nprobe_q,
                        uint32_t nprobe_u, int do_rerank, int metric, const uint32_t* mask,
                        GpuSearchResult* out, uint32_t out_stride, void* workspace,
                        const IvfPqPlan& p, hipStream_t stream);

}  // namespace fv
