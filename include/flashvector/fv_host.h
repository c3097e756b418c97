/*
 * Host search engine: the C++ counterpart of the reference's GpuVectorIndex
 * search orchestration (crates/engine/src/index.rs:1264-1487, 1979-2237;
 * crates/engine/src/streams.rs:22-143), built on the entry points of
 * cuda_bridge.h and fv_ext.h.
 *
 * An engine owns one main index (flat FP16/BF16/FP8 tables with FP32 rows for
 * the rerank, or IVF-PQ tables), a delta buffer of recently added rows
 * (scanned exactly, concurrently with the main search, on a second stream),
 * tombstones, and a pool of streams, each with pre-allocated device scratch
 * (queries, results, masks, workspaces) and pinned host staging, so that a
 * search allocates nothing. Differences from the reference: the slot mask is
 * built on the device from the caller's logical-id words, the tombstones and
 * the slot -> logical-id map (the reference builds it on the host per call);
 * the main + delta merge runs on the device; the device-pointer search makes
 * no host copies and synchronizes only its stream.
 *
 * Thread safety: fvh_search* may be called concurrently from many threads
 * (each call locks one pooled stream's scratch); fvh_load_*, fvh_add_delta
 * and fvh_delete must not run concurrently with searches.
 *
 * Ids: rows are addressed by logical ids (uint32). Results carry logical ids
 * and are ordered by (distance, id); empty slots are {0xFFFFFFFF, 1e30f}.
 * Masks are LSB-first u32 words over logical ids (bit id&31 of word id>>5);
 * words past mask_words count as zero (the reference's unwrap_or(0)).
 * Errors: 0 on success, -1 for a NULL required pointer or a missing index,
 * otherwise a HIP error code or hipErrorInvalidValue for bad shapes.
 */
#ifndef FLASHVECTOR_AMD_FV_HOST_H
#define FLASHVECTOR_AMD_FV_HOST_H

#include "fv_ext.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct FvEngine FvEngine;

typedef struct {
    uint32_t dim;
    uint32_t metric;        /* MetricType */
    uint32_t num_streams;   /* 0 = 4 */
    uint32_t max_batch;     /* queries per device pass; larger calls are split (0 = 1024) */
    uint32_t max_top_k;     /* <= 1024 (0 = 256) */
    uint32_t shortlist_factor; /* flat: shortlist = factor * k (0 = 8) */
    uint32_t flags;         /* FV_SEARCH_* bits for the flat scan */
    uint32_t ivf_nprobe;    /* IVF-PQ: probes per query (0 = 32) */
    uint32_t ivf_shortlist; /* IVF-PQ: rerank depth (0 = 1000) */
    uint32_t reserved;
} FvEngineConfig;

/* crates/engine/src/streams.rs:22-143 (pool + scratchpads). */
int fvh_create(const FvEngineConfig* cfg, FvEngine** out);
void fvh_destroy(FvEngine* e);

/* Main flat index from host rows [n * dim]: FP32 rows uploaded, codes of
 * `dtype` (FvDtype 1..3) and norms encoded on the device. logical_ids (n, or
 * NULL = identity) maps slot -> logical id (index.rs:405-458 id remapping). */
int fvh_load_flat(FvEngine* e, const float* h_vectors, uint32_t n, uint32_t dtype,
                  const uint32_t* h_logical_ids);

/* Main IVF-PQ index from host tables (layout of IvfPqGpuTables; vec_ids are
 * slots) plus optional slot -> logical ids. Codes are also stored in
 * posting-list order for streaming. */
int fvh_load_ivfpq(FvEngine* e, const float* h_centroids, uint32_t nlist,
                   const float* h_codebooks, uint32_t m_pq, const uint8_t* h_codes,
                   const uint32_t* h_offsets, const uint32_t* h_vec_ids, const float* h_vectors,
                   uint32_t n, const uint32_t* h_logical_ids);

/* Appends rows to the delta buffer (exact FP32 scan, merged into every
 * search result). */
int fvh_add_delta(FvEngine* e, const float* h_vectors, const uint32_t* h_logical_ids, uint32_t n);

/* Tombstones logical ids in the main index and the delta buffer. */
int fvh_delete(FvEngine* e, const uint32_t* h_logical_ids, uint32_t n);

/* Host queries [nq * dim] -> host results [nq * top_k]
 * (index.rs:2180-2237 batch_search_fp16_tc, :1979-2108 batch_search_ivf_mode). */
int fvh_search(FvEngine* e, const float* h_queries, uint32_t nq, uint32_t top_k,
               const uint32_t* h_mask_words, uint32_t mask_words, GpuSearchResult* h_out);

/* Device queries -> device results on the caller's stream (the DLPack path
 * of crates/python/src/dlpack.rs:15-121 and index.rs:2110-2178, without the
 * reference's device-wide synchronization). The mask words are device memory
 * (or NULL). */
int fvh_search_device(FvEngine* e, const float* d_queries, uint32_t nq, uint32_t top_k,
                      const uint32_t* d_mask_words, uint32_t mask_words, GpuSearchResult* d_out,
                      void* stream);

/* Queries served and summed wall time of fvh_search* calls (metrics.rs). */
int fvh_stats(const FvEngine* e, uint64_t* queries, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif /* FLASHVECTOR_AMD_FV_HOST_H */
