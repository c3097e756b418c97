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

/* crates/engine/src/index.rs:405-458 build: uploads FP32 rows [n * dim] and
 * builds, on the device, flat tables of `flat_dtype` (FvDtype 1..3, 0 = none)
 * and IVF-PQ tables with min(nlist, n) lists and m_pq sub-quantizers
 * (nlist = 0 = none) by fv_ivf_pq_build (coarse_iters / pq_iters Lloyd
 * iterations, reference 25 / 25; deterministic in `seed`). Identity logical
 * ids; the next fvh_add_vectors id is n. One index may hold both kinds of
 * tables; each search picks one (FvSearchOptions.algo). */
int fvh_build(FvEngine* e, const float* h_vectors, uint32_t n, uint32_t flat_dtype,
              uint32_t nlist, uint32_t m_pq, uint32_t coarse_iters, uint32_t pq_iters,
              uint64_t seed);

/* crates/engine/src/index.rs:716-752 compact: rebuilds the index from its
 * live rows (main rows in slot order, then delta rows) under their logical
 * ids, with the tables it holds (same flat dtype; IVF-PQ with the last
 * build's nlist, iterations and seed, or the loaded tables' nlist); the
 * delta buffer and the tombstones are emptied and the id counter is kept.
 * A no-op without tombstones or delta rows. Synchronizes the device. */
int fvh_compact(FvEngine* e);

/* Appends rows with caller-chosen logical ids to the delta buffer (exact
 * FP32 scan, merged into every search result). Logical ids are unique over
 * the main index, the delta buffer and the deleted ids: a used id, or one
 * repeated in the batch, fails the whole call with hipErrorInvalidValue and
 * changes nothing. Rows deleted earlier stay deleted. */
int fvh_add_delta(FvEngine* e, const float* h_vectors, const uint32_t* h_logical_ids, uint32_t n);

/* crates/engine/src/index.rs:647-694 add_vectors: appends n rows to the delta
 * buffer under fresh logical ids next_logical_id, +1, ..., written to
 * h_assigned_ids[n]. */
int fvh_add_vectors(FvEngine* e, const float* h_vectors, uint32_t n, uint32_t* h_assigned_ids);

/* crates/engine/src/index.rs:583-640 delete_ids: tombstones logical ids held
 * by the main index or the delta buffer (ids already deleted, or held by no
 * row, are skipped). A deleted id stays deleted until a new main index is
 * loaded. */
int fvh_delete(FvEngine* e, const uint32_t* h_logical_ids, uint32_t n);

/* index.rs:2545-2547: the next logical id fvh_add_vectors hands out (one past
 * every id the engine holds, or the loaded snapshot's counter if higher). */
int fvh_next_logical_id(const FvEngine* e, uint32_t* out);

/* Host queries [nq * dim] -> host results [nq * top_k]
 * (index.rs:2180-2237 batch_search_fp16_tc, :1979-2108 batch_search_ivf_mode),
 * with the engine's default options. */
int fvh_search(FvEngine* e, const float* h_queries, uint32_t nq, uint32_t top_k,
               const uint32_t* h_mask_words, uint32_t mask_words, GpuSearchResult* h_out);

/* Per-call options (crates/python/src/py_index.rs:361-436 search_ivf /
 * search_fp16 arguments). algo: which tables serve the call (DEFAULT = flat
 * if loaded, else IVF-PQ; -1 if the engine holds no such tables). nprobe:
 * IVF-PQ probes, 0 = the engine's ivf_nprobe, clamped to nlist. rerank_k:
 * IVF-PQ shortlist with the reference's meaning (ivf_pq_lookup.cu:351-362:
 * 0 = k, 1 = ADC order without rerank, else max(k, rerank_k), at most 10240);
 * negative = the engine's ivf_shortlist with rerank. shortlist_factor: flat
 * shortlist = factor * k, 0 = the engine's. A call shape beyond the engine's
 * configuration grows that stream's workspace first. */
#define FVH_ALGO_DEFAULT 0u
#define FVH_ALGO_FLAT 1u
#define FVH_ALGO_IVFPQ 2u
typedef struct {
    uint32_t algo;
    uint32_t nprobe;
    int32_t rerank_k;
    uint32_t shortlist_factor;
} FvSearchOptions;

int fvh_search_ex(FvEngine* e, const float* h_queries, uint32_t nq, uint32_t top_k,
                  const FvSearchOptions* opts, const uint32_t* h_mask_words, uint32_t mask_words,
                  GpuSearchResult* h_out);

/* Device queries -> device results on the caller's stream (the DLPack path
 * of crates/python/src/dlpack.rs:15-121 and index.rs:2110-2178, without the
 * reference's device-wide synchronization). The mask words are device memory
 * (or NULL). */
int fvh_search_device(FvEngine* e, const float* d_queries, uint32_t nq, uint32_t top_k,
                      const uint32_t* d_mask_words, uint32_t mask_words, GpuSearchResult* d_out,
                      void* stream);

int fvh_search_device_ex(FvEngine* e, const float* d_queries, uint32_t nq, uint32_t top_k,
                         const FvSearchOptions* opts, const uint32_t* d_mask_words,
                         uint32_t mask_words, GpuSearchResult* d_out, void* stream);

/* Engine state (index.rs:2520-2547 num_vectors / delta_len / tombstone_count
 * / next_logical_id). tables: bit 1 flat, bit 2 IVF-PQ. */
typedef struct {
    uint32_t tables;
    uint32_t num_vectors;     /* rows of the main index */
    uint32_t delta_len;       /* rows of the delta buffer, deleted ones included */
    uint32_t tombstones;      /* deleted logical ids */
    uint32_t next_logical_id;
    uint32_t nlist, m_pq;     /* IVF-PQ tables, or 0 */
    uint32_t flat_dtype;      /* FvDtype of the flat tables, or 0 */
} FvEngineInfo;
int fvh_info(const FvEngine* e, FvEngineInfo* out);

/* Queries served and summed wall time of fvh_search* calls (metrics.rs). */
int fvh_stats(const FvEngine* e, uint64_t* queries, double* total_ms);

/* ---- FLSHVEC1 snapshots ---------------------------------------------------
 * The reference's on-disk format (crates/engine/src/persistence.rs:3-200):
 * magic "FLSHVEC1", little-endian u64 manifest length H, H bytes of compact
 * JSON (SnapshotManifest: version, config {dim, max_elements, m,
 * ef_construction, nlist, m_pq, metric}, num_vectors, entry_point,
 * cagra_degree, cagra_seeds, rabitq_padded_dim, fp16_padded_dim,
 * sparse_num_docs, sparse_nnz, next_logical_id, sections {name: {offset,
 * byte_len, dtype, len}} sorted by name), zeros to a 64-byte boundary, then
 * the sections in the order given, each starting 64-byte aligned (offsets
 * relative to that boundary, gaps zero-filled). Host-only; errors: -1 NULL
 * argument, FVH_ERR_IO file error, FVH_ERR_FORMAT invalid snapshot. */
#define FVH_ERR_IO (-2)
#define FVH_ERR_FORMAT (-3)

typedef struct {
    uint32_t version;
    uint32_t dim, max_elements, m, ef_construction, nlist, m_pq, metric; /* IndexConfig */
    uint64_t num_vectors;
    uint32_t entry_point;
    uint32_t next_logical_id;
    uint64_t cagra_degree, rabitq_padded_dim, fp16_padded_dim, sparse_num_docs, sparse_nnz;
    uint32_t num_cagra_seeds; /* written as []; reported when reading */
    uint32_t reserved;
} FvSnapshotInfo;

typedef struct {
    const char* name;    /* section name, e.g. "ivf_centroids" */
    const char* dtype;   /* "f32", "u32", "u16" or "u8" */
    uint64_t len;        /* elements */
    const void* data;    /* len * element-size bytes (may be NULL when len is 0) */
} FvSnapshotSection;

typedef struct FvSnapshot FvSnapshot;

/* write_aligned_snapshot (persistence.rs:51-109). */
int fvh_snapshot_write(const char* path, const FvSnapshotInfo* info,
                       const FvSnapshotSection* sections, uint32_t num_sections);
/* MmapSnapshot::open (persistence.rs:127-162): maps the file read-only and
 * checks magic, manifest bounds and JSON, and that every section lies inside
 * the file. */
int fvh_snapshot_open(const char* path, FvSnapshot** out, FvSnapshotInfo* info);
/* Section bytes inside the mapping (valid until fvh_snapshot_close); returns 1
 * when the section is absent. dtype receives up to 7 characters. */
int fvh_snapshot_section(const FvSnapshot* s, const char* name, const void** data,
                         uint64_t* byte_len, uint64_t* len, char* dtype8);
void fvh_snapshot_close(FvSnapshot* s);

/* ---- Micro-batcher ---------------------------------------------------------
 * Counterpart of the reference server's SearchBatcher
 * (crates/server/src/batcher.rs:80-262): callers submit one query at a time
 * from any number of threads; a worker thread takes the first waiting query,
 * keeps collecting until max_batch queries are held or max_wait_us has passed
 * since the first, splits what it holds by launch key (top_k here: the
 * engine fixes metric, nprobe and shortlist; groups keep the order of their
 * first query), runs each group as one batched search and hands every caller
 * its own k results. A failed batch returns its error code to every query of
 * the group. fvh_batcher_destroy serves the queries already queued, then
 * stops the worker; it must not run concurrently with fvh_batcher_submit.
 * Errors: -1 NULL argument, hipErrorInvalidValue for top_k 0,
 * FVH_ERR_SHUTDOWN for a submit after destroy began, else the executor's. */
#define FVH_ERR_SHUTDOWN (-4)

typedef struct FvBatcher FvBatcher;

/* Runs num_queries queries ([num_queries * dim] f32, host) at top_k and
 * writes num_queries * top_k results (host); returns 0 or an error code. */
typedef int (*FvBatchExecutor)(void* user, uint32_t top_k, const float* queries,
                               uint32_t num_queries, GpuSearchResult* out);

/* SearchBatcher::for_index (batcher.rs:89-124; the reference server uses
 * max_wait 350 us, max_batch 64): the executor is fvh_search on `e`, which
 * must outlive the batcher. max_batch 0 counts as 1. */
int fvh_batcher_create(FvEngine* e, uint32_t max_wait_us, uint32_t max_batch, FvBatcher** out);

/* SearchBatcher::with_executor (batcher.rs:126-229) for any executor of
 * queries of `dim` floats. */
int fvh_batcher_create_with_executor(FvBatchExecutor exec, void* user, uint32_t dim,
                                     uint32_t max_wait_us, uint32_t max_batch, FvBatcher** out);

/* SearchBatcher::submit (batcher.rs:231-247): blocks until the query's batch
 * ran. h_query holds dim floats, h_out receives top_k results; latency_us
 * (may be NULL) receives the batch's executor time in microseconds. */
int fvh_batcher_submit(FvBatcher* b, const float* h_query, uint32_t top_k,
                       GpuSearchResult* h_out, double* latency_us);

/* SearchBatcher::stats (batcher.rs:249-262); avg_batch_size is 0 before the
 * first batch. Any output may be NULL. */
int fvh_batcher_stats(const FvBatcher* b, uint64_t* total_requests, uint64_t* total_batches,
                      double* avg_batch_size);

void fvh_batcher_destroy(FvBatcher* b);

/* Engine state as a snapshot (index.rs:845-1031): FP32 rows, logical ids,
 * slot validity words, IVF-PQ tables, FP16 rows padded to 16 and their
 * norms (flat FP16 engines; BF16 / FP8 codes go to the extra sections
 * "vectors_bf16" / "vectors_fp8" + "row_scales"), tombstoned logical ids,
 * delta rows and ids. Graph, SQ8 and RaBitQ sections are written empty. */
int fvh_save_snapshot(FvEngine* e, const char* path);
/* Loads a snapshot into an engine created with the same dim and metric
 * (index.rs:1033-1258 for the sections above): flat_dtype 0 loads the IVF-PQ
 * tables (which must be present), FvDtype 1..3 a flat index encoded from the
 * FP32 rows, FvDtype | FVH_LOAD_IVFPQ both; then the delta rows, then the
 * tombstones (or, without a tombstones section, the cleared bits of
 * main_valid_words). */
#define FVH_LOAD_IVFPQ 0x100u
int fvh_load_snapshot(FvEngine* e, const char* path, uint32_t flat_dtype);

#ifdef __cplusplus
}
#endif
#endif /* FLASHVECTOR_AMD_FV_HOST_H */
