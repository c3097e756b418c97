/*
 * IVF-PQ index tables built on the GPU: the step before the query scan.
 *
 * Replaces the CPU build of the reference's IVF-PQ tables
 * (crates/engine/src/index.rs:405-458 with KMeans::fit,
 * crates/engine/src/cluster.rs:28-182, and ProductQuantizer::train / encode,
 * cluster.rs:197-287). None of these entry points exist in the reference, whose
 * build runs inside GpuVectorIndex::build on the host.
 *
 * Algorithm: the reference's, made deterministic. Every random draw comes from
 * splitmix64(seed) in a fixed order (a partial Fisher-Yates shuffle choosing
 * the k-means++ pool of min(n, 64k) rows, k-1 seeding targets, then
 * empty-cluster recovery draws in ascending cluster order), and the k-means++
 * target is located by f64 sums over 256-candidate chunks. FP32 arithmetic is
 * the reference's: l2 = sequential non-fused sum of (x - c)^2; assignment =
 * first strict minimum; centroid = sum of its rows in row order times
 * (1.0f / count); iterations stop when the largest squared centroid move is
 * below 1e-4f. Given the same initial centroids and no empty cluster, the
 * Lloyd iterations reproduce the reference's bit for bit.
 *
 * These calls are synchronous: they read cluster sizes back every iteration
 * and return after the work on `stream` has finished. Workspaces follow the
 * fv_ext.h convention (device memory, 256-byte aligned, at least the size the
 * *_workspace_size call returns, not shared between concurrent calls).
 */
#ifndef FLASHVECTOR_AMD_FV_BUILD_H
#define FLASHVECTOR_AMD_FV_BUILD_H

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Nearest of k centroids ([k * dim], row-major) for every row x + i*ld
 * (ld >= dim): the assign step of cluster.rs:97-123 and the encode search of
 * cluster.rs:256-279. d_assign [n] (may be NULL) gets the index; d_codes, when
 * non-NULL, gets it as a byte at d_codes[i * code_stride] (k <= 256). */
int fv_kmeans_assign(const float* d_x, uint32_t n, uint32_t dim, uint32_t ld,
                     const float* d_centroids, uint32_t k, uint32_t* d_assign,
                     uint8_t* d_codes, uint32_t code_stride, void* stream);

/* KMeans::fit (cluster.rs:28-182), deterministic as above, over rows
 * x + i*ld. Outputs: d_centroids [k * dim]; d_assign [n], the last assign
 * step's clusters (the reference's `assignments`); d_sizes [k] their counts;
 * d_offsets [k + 1] / d_members [n] (may be NULL) the rows of each cluster
 * ascending; *iters_run (host, may be NULL) the Lloyd iterations done.
 * Requires 1 <= k <= n. */
size_t fv_kmeans_workspace_size(uint32_t n, uint32_t dim, uint32_t k);
int fv_kmeans_fit(const float* d_x, uint32_t n, uint32_t dim, uint32_t ld, uint32_t k,
                  uint32_t max_iters, uint64_t seed, float* d_centroids, uint32_t* d_assign,
                  uint32_t* d_sizes, uint32_t* d_offsets, uint32_t* d_members,
                  uint32_t* iters_run, void* d_workspace, size_t workspace_bytes, void* stream);

/* The IVF-PQ tables of index.rs:405-458 from FP32 rows [n * dim]:
 * nlist' = min(nlist, n) (at least 1) coarse centroids from k-means over all
 * rows (coarse_iters, reference 25), residuals against the assigned
 * centroid, one k-means per sub-space of the residuals with min(n, 256)
 * codewords (pq_iters, reference 25; sub-quantizer m seeded with
 * seed + (m + 1) * 0x9E3779B97F4A7C15; unused codewords FLT_MAX), residual
 * codes [n * m_pq] by row, posting lists ascending by row.
 * Outputs: d_centroids [nlist' * dim], d_codebooks [m_pq * 256 * (dim/m_pq)],
 * d_codes [n * m_pq], d_ivf_offsets [nlist' + 1], d_ivf_vec_ids [n];
 * iters_run (host, may be NULL) [1 + m_pq]. dim % m_pq == 0. */
size_t fv_ivf_pq_build_workspace_size(uint32_t n, uint32_t dim, uint32_t nlist, uint32_t m_pq);
int fv_ivf_pq_build(const float* d_vectors, uint32_t n, uint32_t dim, uint32_t nlist,
                    uint32_t m_pq, uint32_t coarse_iters, uint32_t pq_iters, uint64_t seed,
                    float* d_centroids, float* d_codebooks, uint8_t* d_codes,
                    uint32_t* d_ivf_offsets, uint32_t* d_ivf_vec_ids, uint32_t* iters_run,
                    void* d_workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FLASHVECTOR_AMD_FV_BUILD_H */
