/*
 * CPU oracle for the FlashVector query-scan hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates, in plain C, the arithmetic
 * and selection contract of the reference's CPU path and of its scan kernels,
 * so that the HIP implementation can be checked against it on identical
 * inputs. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as the checker or as the timed CPU baseline; the
 * product library never links or calls it.
 *
 * Selection contract everywhere: the k smallest by (distance, id), ties going
 * to the lower id (the stable sort of crates/engine/src/metrics.rs:125-129 and
 * the strict-< insertion of kernels/src/sparse_hybrid.cu:60-73). Only
 * candidates with distance < 1e30f are admitted (the sentinel the reference
 * kernels initialise their lists with). Empty result slots are {0xFFFFFFFF, 1e30f}.
 *
 * Exact FP32 distance: a sequential, non-fused sum over d = 0..dim-1, in the
 * order of the reference CPU path (metrics.rs:112-121: diff = q - v; acc +=
 * diff * diff; no FMA contraction). Inner product and cosine use the same
 * sequential non-fused order for their dot and norm sums with the reference's
 * definitions (kernels/src/distance_metrics.cuh:71-118): IP = -dot, cosine =
 * 1 - dot / (sqrtf(|q|^2) * sqrtf(|x|^2) + 1e-8f).
 *
 * Parity pins (tests/test_oracle.py): closed-form sums of the reference's
 * distance KAT (tests/cuda_sanity_test.rs:44-113), FP64 brute force on integer
 * data where FP32 is exact, the reference's bitonic KAT
 * (tests/cuda_sanity_test.rs:115-159), FP16 round-trip values
 * (crates/engine/src/quant.rs:362-368) plus an exhaustive binary16 sweep
 * against numpy, and IVF-PQ invariants (exhaustive probing equals the flat
 * scan; integer codebooks make ADC equal a decode-and-brute-force).
 * FP8 e4m3 conversion is pinned exhaustively against torch's float8_e4m3fn
 * and a brute-force nearest-value search; the FP64 top-k against numpy.
 * Functions without such a check are marked "parity unpinned" below.
 */
#ifndef FV_ORACLE_H
#define FV_ORACLE_H

#include <stdint.h>

#include "../include/flashvector/types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Exact FP32 distance between q and x (sequential, non-fused).
 * L2: crates/engine/src/metrics.rs:112-121. IP/cosine definitions:
 * kernels/src/distance_metrics.cuh:71-118. */
float fvo_distance(const float* q, const float* x, uint32_t dim, int metric);

/* RecallEvaluator::exact_knn (crates/engine/src/metrics.rs:103-135): L2 over
 * all n rows, stable ascending order, truncated to top_k. Writes min(top_k, n)
 * entries and returns that count. */
uint32_t fvo_exact_knn(const float* dataset, uint32_t n, uint32_t dim,
                       const float* query, uint32_t top_k, GpuSearchResult* out);

/* cuda_flat_scan_batch contract (kernels/src/sparse_hybrid.cu:13-107,504-524):
 * Q==0 or top_k==0 -> nothing written; top_k is clamped to 256 and is also the
 * output row stride; n==0 -> every output byte 0xFF; the mask is tested on the
 * slot index; ids are remapped through vec_ids when non-NULL; ties go to the
 * lower slot. Returns 0, or -1 for a NULL required pointer. */
int fvo_flat_scan_batch(const float* vectors, const uint32_t* vec_ids, uint32_t n,
                        uint32_t dim, const float* queries, uint32_t nq, uint32_t top_k,
                        int metric, const uint32_t* valid_mask, GpuSearchResult* out,
                        int nthreads);

/* Same selection without the 256 clamp (row stride = top_k): the contract of
 * the extended flat entry points when they rerank exactly in FP32. */
int fvo_flat_topk(const float* vectors, uint32_t n, uint32_t dim, const float* queries,
                  uint32_t nq, uint32_t top_k, int metric, const uint32_t* valid_mask,
                  GpuSearchResult* out, int nthreads);

/* Coarse quantizer of kernels/src/ivf_rerank.cuh:49-130: per centroid an fmaf
 * chain over d of (q - c)^2, then the nprobe smallest by (dist, cluster).
 * Writes nprobe ids (0xFFFFFFFF when fewer centroids qualify). */
void fvo_coarse_probe(const float* query, const float* centroids, uint32_t dim,
                      uint32_t nlist, uint32_t nprobe, uint32_t* probed, float* probed_dist);

/* Residual LUT of kernels/src/ivf_pq_lookup.cu:150-162: lut[m*256 + j] = fmaf
 * chain over s < sub_dim of ((q - c) - B[m][j])^2 (always L2). */
void fvo_pq_lut(const float* query, const float* centroid, const float* codebooks,
                uint32_t m_pq, uint32_t sub_dim, float* lut);

/* cuda_ivf_pq_search_batch contract (kernels/src/ivf_pq_lookup.cu:339-473)
 * with the deterministic (distance, id) ties rule: clamps top_k <= 256,
 * nprobe to [1, nlist], shortlist S = (rerank_k ? rerank_k : top_k) in
 * [top_k, 10240]; ADC = sequential FP32 adds of LUT entries m = 0..M-1
 * (:176-178); rerank with fvo_distance when rerank_k != 1 and vectors != NULL.
 * Tables hold host pointers. */
int fvo_ivf_pq_search_batch(const IvfPqGpuTables* t, const float* queries, uint32_t nq,
                            uint32_t top_k, uint32_t nprobe, uint32_t rerank_k, int metric,
                            const uint32_t* valid_mask, GpuSearchResult* out, int nthreads);

/* Uneven batch: per-query k[i] (<= 1024) and nprobe[i]; shortlist depth is
 * max(shortlist, k[i]) capped at 16384; rows of out have stride out_stride and
 * are padded past k[i]. No 256 clamp. */
int fvo_ivf_pq_search_uneven(const IvfPqGpuTables* t, const float* queries, uint32_t nq,
                             const uint32_t* k_per_query, const uint32_t* nprobe_per_query,
                             uint32_t shortlist, int rerank, int metric,
                             const uint32_t* valid_mask, GpuSearchResult* out,
                             uint32_t out_stride, int nthreads);

/* cuda_merge_topk contract (kernels/src/sparse_hybrid.cu:139-218) with the
 * host_merge_topk rule (crates/engine/src/index.rs:2637-2678): per query the
 * first min(num_shards*shard_top_k, 4096) shard entries (shard-major), empty
 * ids dropped, ordered by (distance, id), first occurrence of each id kept,
 * out_top_k clamped to 256, padded. */
int fvo_merge_topk(const GpuSearchResult* shard_results, uint32_t num_shards,
                   uint32_t nq, uint32_t shard_top_k, uint32_t out_top_k,
                   GpuSearchResult* out);

/* binary16 conversions of crates/engine/src/quant.rs:239-304 (round half up
 * on magnitude, subnormals truncated). */
uint16_t fvo_f32_to_f16_bits(float x);
float fvo_f16_bits_to_f32(uint16_t h);

/* encode_fp16_padded (crates/engine/src/quant.rs:306-327): rows padded to a
 * multiple of 16, norms are sequential FP32 sums of the FP32 originals.
 * Returns padded_dim. */
uint32_t fvo_encode_fp16_padded(const float* vectors, uint32_t n, uint32_t dim,
                                uint16_t* out_fp16, float* out_norms);

/* compute_distances_warp KAT reference (kernels/src/hnsw_traverse.cu:387-415):
 * the dense [nq x n] matrix of fvo_distance. */
void fvo_distance_matrix(const float* queries, const float* dataset, uint32_t nq,
                         uint32_t n, uint32_t dim, int metric, float* out);

/* OCP FP8 E4M3 (bias 7, no infinities, largest finite 448): round to nearest
 * even, saturating to +-448 (NaN -> 0x7f), and its exact decode. The format
 * is the one of v_mfma_f32_16x16x128_f8f6f4 with format code 0. */
uint8_t fvo_f32_to_e4m3(float x);
float fvo_e4m3_to_f32(uint8_t c);

/* FP8 rows with one FP32 scale per row: s = max|x| / 448 (1 for an all-zero
 * row), code = e4m3(sign * x / s), zero padded to pdim codes. */
void fvo_encode_fp8_rows(const float* x, uint32_t n, uint32_t dim, uint32_t pdim, float sign,
                         uint8_t* codes, float* scales);

/* Exact top-k with distances evaluated in FP64 (float products are exact in
 * FP64; L2 = sum (q-x)^2, IP = -sum q*x, cosine = 1 - dot/(|q||x| + 1e-8)),
 * ties to the lower id. The FP64 oracle of recall and distance-error reports
 * and of the FP64-boundary search mode. */
int fvo_flat_topk_f64(const float* vectors, uint32_t n, uint32_t dim, const float* queries,
                      uint32_t nq, uint32_t top_k, int metric, const uint32_t* valid_mask,
                      uint32_t* out_ids, double* out_dist, int nthreads);

/* ---- IVF-PQ index construction -------------------------------------------
 * Deterministic restatement of KMeans::fit (crates/engine/src/cluster.rs:28-182),
 * ProductQuantizer::train/encode (cluster.rs:197-287) and the IVF-PQ part of
 * GpuVectorIndex::build (crates/engine/src/index.rs:405-458). The reference
 * draws from an unseeded thread RNG; here every draw comes from splitmix64 in
 * a fixed order (pool shuffle of min(n, 64k) candidates, k-1 seeding targets,
 * then empty-cluster recovery in ascending cluster order) and the k-means++
 * target is found by f64 sums over 256-candidate chunks (first candidate
 * whose running sum exceeds u * total). FP32 steps follow the reference:
 * l2 = sequential non-fused sum of (x - c)^2; assignment = first strict
 * minimum (best starts at +inf, cluster 0 if none); centroid = sum in
 * vector-index order times (1.0f / count); stop when the largest squared move
 * is below 1e-4f. Rows are x + i*ld (ld >= dim). Parity pins: closed forms
 * (integer data with power-of-two cluster sizes, separated blobs, codeword
 * identity encode) and brute force in tests/test_oracle.py; the seeding
 * search order is pinned only through invariants (parity partly unpinned). */
void fvo_kmeans_assign(const float* x, uint32_t n, uint32_t dim, uint32_t ld,
                       const float* centroids, uint32_t k, uint32_t* assign);
void fvo_kmeans_fit(const float* x, uint32_t n, uint32_t dim, uint32_t ld, uint32_t k,
                    uint32_t max_iters, uint64_t seed, float* centroids, uint32_t* assign,
                    uint32_t* sizes, uint32_t* iters_run);
int fvo_ivf_pq_build(const float* x, uint32_t n, uint32_t dim, uint32_t nlist, uint32_t m_pq,
                     uint32_t coarse_iters, uint32_t pq_iters, uint64_t seed, float* centroids,
                     float* codebooks, uint8_t* codes, uint32_t* offsets, uint32_t* vec_ids,
                     uint32_t* iters_run);

#ifdef __cplusplus
}
#endif
#endif /* FV_ORACLE_H */
