/*
 * Additive entry points of the MI355X query-scan library. None of these exist
 * in the reference; they lift its host limits (top_k <= 256, per-call device
 * allocation) and expose the reduced-precision scan, uneven batches and the
 * top-k backward pass. The reference-named entry points in cuda_bridge.h are
 * unchanged by them.
 *
 * Conventions as in cuda_bridge.h. Entry points that take a workspace never
 * allocate: the caller passes at least *_workspace_size() bytes of device
 * memory (256-byte aligned) and must not share it between concurrent calls.
 */
#ifndef FLASHVECTOR_AMD_FV_EXT_H
#define FLASHVECTOR_AMD_FV_EXT_H

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FV_MAX_TOPK 1024u
#define FV_MAX_SHORTLIST 16384u

/* ---- exact FP32 flat scan that allocates nothing -----------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cuda_flat_scan_batch's arguments, clamps and results,
 * with its scratch in a static 64 MiB array per GPU held by the library until
 * the process ends. Calls that overlap on one GPU share that array and race;
 * serialize them. Returns hipErrorOutOfMemory, launching nothing, when the
 * scratch of one row chunk does not fit: about
 * num_queries * (next_pow2(k + 128) + k) * 8 bytes (k = min(top_k, 256)),
 * plus (num_vectors + num_queries) * 4 for cosine. */
int fv_flat_scan_batch_noalloc(const float* d_vectors, const uint32_t* d_vec_ids,
                               uint32_t num_vectors, uint32_t dim, const float* d_queries,
                               uint32_t num_queries, uint32_t top_k, MetricType metric,
                               const uint32_t* d_valid_mask, GpuSearchResult* d_out_results,
                               void* stream);

/* ---- exact FP32 flat scan without the 256 clamp ------------------------ */
size_t fv_flat_exact_workspace_size(uint32_t num_vectors, uint32_t dim, uint32_t num_queries,
                                    uint32_t top_k, MetricType metric);
int fv_flat_exact_search(const float* d_vectors, uint32_t num_vectors, uint32_t dim,
                         const float* d_queries, uint32_t num_queries, uint32_t top_k,
                         MetricType metric, const uint32_t* d_valid_mask,
                         GpuSearchResult* d_out_results, void* d_workspace,
                         size_t workspace_bytes, void* stream);

/* ---- reduced-precision flat scan on the matrix cores --------------------
 * Rows are stored as 16-bit codes (FP16 by the reference's host conversion,
 * crates/engine/src/quant.rs:237-273, i.e. halves rounded away from zero and
 * subnormals truncated; BF16 round-to-nearest-even) or as FP8
 * e4m3 codes with one FP32 scale per row (s = max|x|/448, code =
 * RNE-saturating e4m3(x/s)), padded with zeros to padded_dim (a multiple of
 * 16, at most 768). The scan scores every row on MFMA with FP32 accumulation
 * (queries are encoded the same way per call, FP8 with one scale per query),
 * keeps a shortlist per query, reranks it exactly in FP32 against d_vectors
 * (the reference CPU path's arithmetic) and returns the top_k by (distance,
 * row). With d_vectors NULL the reduced-precision scores are returned. */
typedef enum {
    FV_DTYPE_F32 = 0,
    FV_DTYPE_F16 = 1,
    FV_DTYPE_BF16 = 2,
    FV_DTYPE_FP8_E4M3 = 3
} FvDtype;

typedef struct {
    const void* d_codes;          /* [num_vectors * padded_dim] codes of `dtype` */
    const float* d_row_scales;    /* [num_vectors] dequantisation scale (FP8) or NULL */
    const float* d_norms_sq;      /* [num_vectors] |x|^2 of the FP32 rows */
    const double* d_norms_sq_f64; /* [num_vectors] |x|^2 in FP64, or NULL */
    const float* d_vectors;       /* [num_vectors * dim] FP32 rows for the rerank, or NULL */
    uint32_t num_vectors;
    uint32_t dim;
    uint32_t padded_dim;
    uint32_t dtype;               /* FvDtype */
    const float* d_enc_err;       /* [num_vectors] |decode(code) - x| of FP16/BF16 rows
                                     (fv_encode_error), or NULL (a bound from |x| is used) */
} FvFlatTables;

/* Encodes FP32 rows into FvFlatTables storage on the device: codes (padded
 * with zeros; FP8 also writes d_row_scales_out), FP32 norms as sequential
 * non-fused sums, FP64 norms. Any output pointer may be NULL to skip it. */
int fv_encode_flat(const float* d_vectors, uint32_t num_vectors, uint32_t dim,
                   uint32_t padded_dim, uint32_t dtype, void* d_codes_out,
                   float* d_norms_out, double* d_norms_f64_out, float* d_row_scales_out,
                   void* stream);
/* Per-row Euclidean norm of the encoding error of FP16/BF16 codes,
 * |decode(code_row) - x_row|, rounded up (FP64 sums), for the bounded rerank
 * of fv_flat_search. Other dtypes: hipErrorInvalidValue. */
int fv_encode_error(const float* d_vectors, const void* d_codes, uint32_t num_vectors, uint32_t dim,
                    uint32_t padded_dim, uint32_t dtype, float* d_err_out, void* stream);

/* Search parameters of the flat scans. Shortlist of query i: factor * k_i
 * when shortlist_factor != 0, else `shortlist` (0 means 8 * k_i), never below
 * k_i; at most 7936 rows. valid_count is the number of set bits in the mask
 * when the caller knows it (0 = unknown): a mask keeping at most half of the
 * rows is then compacted on the device and only the valid rows are scored.
 * FV_VALID_COUNT_DEVICE: the mask is always compacted and the count is read
 * on the device, where the scan sizes its row chunks from it (no host
 * count, no synchronization; the workspace is planned for every row).
 * FV_SEARCH_F64_BOUNDARY (needs d_vectors): reranked candidates whose FP32
 * distances are within twice their rounding-error bound of a neighbour are
 * re-scored in FP64 and ordered by (FP64 distance, row), so membership and
 * order of the returned top k follow the exact distances; those rows report
 * the FP64 distance rounded to FP32.
 * Without it, L2 and inner-product searches over FP16/BF16 codes rerank only
 * the shortlisted rows whose exact-distance interval (scan score +- a
 * rigorous bound from the codes' rounding error, d_enc_err when given, and
 * FP32 accumulation) reaches the k-th smallest upper end: no other row can
 * be among the k best, so the results equal those of reranking every row.
 * FV_SEARCH_FULL_RERANK turns that off (every shortlisted row is reranked). */
#define FV_SEARCH_F64_BOUNDARY 1u
#define FV_SEARCH_FULL_RERANK 2u
#define FV_VALID_COUNT_DEVICE 0xFFFFFFFFu
typedef struct {
    const uint32_t* d_valid_mask; /* LSB-first u32 words, NULL = every row */
    const uint32_t* d_top_k;      /* per-query k (each <= top_k), or NULL */
    uint32_t valid_count;
    uint32_t metric;              /* MetricType */
    uint32_t top_k;               /* uniform k and output row stride, <= 1024 */
    uint32_t shortlist_factor;
    uint32_t shortlist;
    uint32_t flags;               /* FV_SEARCH_* bits, others must be 0 */
    /* Plan overrides (0 = automatic), normally set from fv_flat_autotune:
     * row chunks per query block (the scan's work-item shape), and the
     * fraction of rows the first pass scans as a divisor (1, 2, 4, ...;
     * 0xFFFFFFFF = no first pass). */
    uint32_t n_chunks;
    uint32_t bound_div;
    /* With d_top_k: an upper bound on the sum of the per-query k (0 = not
     * given). Deep shortlists then get per-query candidate buffers sized from
     * each query's shortlist, laid out on the device, instead of buffers of
     * the deepest shortlist for every query (far less workspace when k
     * varies). A query whose buffer does not fit because the sum was
     * understated returns no rows (every slot empty). */
    uint32_t top_k_sum;
} FvFlatSearchParams;

size_t fv_flat_workspace_size(const FvFlatTables* tables, uint32_t num_queries,
                              const FvFlatSearchParams* params);
int fv_flat_search(const FvFlatTables* tables, const float* d_queries, uint32_t num_queries,
                   const FvFlatSearchParams* params, GpuSearchResult* d_out_results,
                   void* d_workspace, size_t workspace_bytes, void* stream);

/* Autotuner of the flat search for one (table, query distribution, k,
 * metric, dtype): on the query sample d_queries it (1) computes the exact
 * FP32 top-k (fv_flat_exact_search, so tables->d_vectors is required), (2)
 * picks the smallest shortlist factor in {1, 2, 4, 8, 16} whose recall@k
 * against it is at least target_recall (16 if none) — or keeps
 * base->shortlist_factor when that is non-zero — (3) times every
 * combination of n_chunks in {auto/2, auto, 2, 4, 8 x auto} and bound_div in
 * {auto, 2*auto, 4*auto, none} with that factor (median of `trials` runs
 * after a warm-up) and returns the fastest, keeping the automatic plan
 * (n_chunks = bound_div = 0) unless a candidate is at least 3% faster. `base` supplies k, metric, mask and
 * flags. Allocates and frees its own device memory and synchronizes the
 * stream: a set-up call, not for the search path. */
typedef struct {
    uint32_t shortlist_factor;
    uint32_t n_chunks;
    uint32_t bound_div;
    uint32_t reserved;
    float ms;                     /* best time of the chosen plan */
    float recall;                 /* recall@k of the chosen shortlist factor */
} FvFlatTuning;
int fv_flat_autotune(const FvFlatTables* tables, const float* d_queries, uint32_t num_queries,
                     const FvFlatSearchParams* base, float target_recall, uint32_t trials,
                     FvFlatTuning* out_best, void* stream);

/* ---- IVF-PQ scan with uneven batches ------------------------------------
 * Same arithmetic and ties rule as cuda_ivf_pq_search_batch, without its 256
 * and 10240 clamps: per-query d_top_k[i] (<= max_top_k <= 1024) and
 * d_nprobe[i] (<= max_nprobe) when those pointers are non-NULL (else the
 * uniform max values apply), shortlist depth max(shortlist, k_i) <= 16384.
 * Output rows have stride max_top_k and are padded past k_i. When
 * d_codes_list_order is non-NULL it must hold the codes in posting-list order
 * (fv_ivf_pq_list_order_codes), which turns the per-entry gathers into
 * coalesced streaming reads. Any nlist. */
int fv_ivf_pq_list_order_codes(const IvfPqGpuTables* tables, uint8_t* d_codes_out, void* stream);
size_t fv_ivf_pq_workspace_size(const IvfPqGpuTables* tables, uint32_t num_queries,
                                uint32_t max_nprobe, uint32_t shortlist, uint32_t max_top_k);
int fv_ivf_pq_search(const IvfPqGpuTables* tables, const uint8_t* d_codes_list_order,
                     const float* d_queries, uint32_t num_queries, const uint32_t* d_top_k,
                     uint32_t max_top_k, const uint32_t* d_nprobe, uint32_t max_nprobe,
                     uint32_t shortlist, int rerank, MetricType metric,
                     const uint32_t* d_valid_mask, GpuSearchResult* d_out_results,
                     void* d_workspace, size_t workspace_bytes, void* stream);

/* Work-item plan of the IVF-PQ scan. The batch is scanned by persistent
 * workgroups over work items: a query whose probed lists hold more than C
 * entries, C = max(chunk_min, total / target_items, 2 * total / slots) with
 * `total` the batch's probed entries and `slots` fixed by the shortlist
 * depth, is cut into items of at most C entries (its lists concatenated in
 * probe order) that several workgroups scan at once; the item that finishes
 * last merges their shortlists. target_items = 0 keeps one item per query.
 * Results do not depend on the plan. */
typedef struct {
    uint32_t target_items;
    uint32_t chunk_min;           /* entries; 0 = library default */
    float ms;                     /* autotuner: median time of the batch with this plan */
    uint32_t reserved;
} FvIvfPqTuning;
/* fv_ivf_pq_search with an explicit work-item plan (NULL = library default). */
int fv_ivf_pq_search_tuned(const IvfPqGpuTables* tables, const uint8_t* d_codes_list_order,
                           const float* d_queries, uint32_t num_queries, const uint32_t* d_top_k,
                           uint32_t max_top_k, const uint32_t* d_nprobe, uint32_t max_nprobe,
                           uint32_t shortlist, int rerank, MetricType metric,
                           const uint32_t* d_valid_mask, GpuSearchResult* d_out_results,
                           void* d_workspace, size_t workspace_bytes,
                           const FvIvfPqTuning* tuning, void* stream);
/* Autotuner of the IVF-PQ work-item plan for one batch shape: times the
 * search (median of `trials` runs after a warm-up) over target_items in
 * {0, 256, 1024, 4096, 16384} x chunk_min in {4096, 16384, 65536} and
 * returns the fastest, keeping the library default unless a candidate is at
 * least 3% faster. Allocates and frees its own device memory and
 * synchronizes the stream: a set-up call, not for the search path. */
int fv_ivf_pq_autotune(const IvfPqGpuTables* tables, const uint8_t* d_codes_list_order,
                       const float* d_queries, uint32_t num_queries, const uint32_t* d_top_k,
                       uint32_t max_top_k, const uint32_t* d_nprobe, uint32_t max_nprobe,
                       uint32_t shortlist, int rerank, MetricType metric,
                       const uint32_t* d_valid_mask, uint32_t trials, FvIvfPqTuning* out_best,
                       void* stream);

/* ---- top-k inner-product backward ----------------------------------------
 * Scores s_ij = q_i . x_{ids[i][j]} of the rows a top-k search selected
 * (d_ids [num_queries * top_k], 0xFFFFFFFF = empty). With d_grad_scores
 * (dL/ds, same shape) the kernel forms dL/dq_i = sum_j g_ij x_{ids[i][j]}
 * into d_grad_queries and, when d_grad_vectors is non-NULL, scatter-adds
 * g_ij q_i into the selected rows of d_grad_vectors ([num_vectors * dim],
 * accumulated with float atomics: the caller zeroes it). With
 * d_grad_scores NULL it recomputes the scores and uses the gradient of the
 * softmax cross-entropy of s_i * inv_tau against the target position
 * d_targets[i], times loss_scale and times *d_grad_output when that device
 * scalar is non-NULL (the upstream gradient of the loss, read on the device);
 * the recomputed scores are written to d_scores_out when non-NULL.
 * dim <= 4096, top_k <= 1024. No counterpart in the reference (it has no
 * training path). */
int fv_topk_ip_backward(const float* d_queries, uint32_t num_queries, const float* d_vectors,
                        uint32_t num_vectors, uint32_t dim, const uint32_t* d_ids,
                        uint32_t top_k, const float* d_grad_scores, const uint32_t* d_targets,
                        float inv_tau, float loss_scale, const float* d_grad_output,
                        float* d_grad_queries, float* d_grad_vectors, float* d_scores_out,
                        void* stream);

/* ---- kernel timer --------------------------------------------------------
 * While enabled, every search call brackets its dominant scan kernel with two
 * HIP events recorded on the call's stream. fv_kernel_timer_read waits for the
 * recorded events and returns their summed elapsed time and count (reset != 0
 * clears them). Process-wide; meant for benchmarks. */
int fv_kernel_timer_enable(int enable);
int fv_kernel_timer_read(double* total_ms, uint32_t* launches, int reset);

#ifdef __cplusplus
}
#endif
#endif /* FLASHVECTOR_AMD_FV_EXT_H */
