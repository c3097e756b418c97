/*
 * Plain-old-data types of the FlashVector query-scan ABI.
 *
 * Every struct here is byte-identical to its counterpart in the reference
 * (kernels/include/types.h); the layouts are pinned by static_asserts below and
 * by tests/test_abi.py, because Rust binds them through #[repr(C)] mirrors
 * (crates/engine/src/ffi.rs:20-61,146-159,200-232).
 */
#ifndef FLASHVECTOR_AMD_TYPES_H
#define FLASHVECTOR_AMD_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#define FV_ALIGN8 alignas(8)
#else
#define FV_ALIGN8 _Alignas(8)
#endif
/* The 8-byte alignment sits on the first member so that one spelling works in
 * C and C++; the struct's alignment is the same as the reference's. */

/* Distance metric (types.h:15-19). L2 is the squared distance, IP is -dot,
 * cosine is 1 - dot / (|q| |x| + 1e-8). Smaller is better for all three. */
typedef enum {
    METRIC_L2 = 0,
    METRIC_COSINE = 1,
    METRIC_INNER_PRODUCT = 2
} MetricType;

/* One result slot (types.h:39-42): 8 bytes, 8-byte aligned. Empty slots hold
 * {0xFFFFFFFF, 1e30f}. */
typedef struct {
    FV_ALIGN8 uint32_t id;
    float distance;
} GpuSearchResult;

/* IVF-PQ tables (types.h:85-97). Codes are indexed by vector id, lists hold
 * vector ids in ascending order per cluster. d_vectors may be NULL (no rerank). */
typedef struct {
    const float* d_centroids;      /* [nlist * dim] */
    const float* d_pq_codebooks;   /* [m_pq * 256 * sub_dim] */
    const uint8_t* d_pq_codes;     /* [num_vectors * m_pq] */
    const uint32_t* d_ivf_offsets; /* [nlist + 1] */
    const uint32_t* d_ivf_vec_ids; /* [num_vectors] */
    const float* d_vectors;        /* [num_vectors * dim] FP32 rows for rerank */
    uint32_t num_vectors;
    uint32_t dim;
    uint32_t nlist;
    uint32_t m_pq;
    uint32_t sub_dim;
} IvfPqGpuTables;

/* FP16 flat tables (types.h:131-138). Rows are padded to a multiple of 16
 * halves; d_vec_norms_sq holds ||x||^2 of the FP32 originals. */
typedef struct {
    const uint16_t* d_vectors_fp16; /* [num_vectors * padded_dim] binary16 bits */
    const float* d_vec_norms_sq;    /* [num_vectors] */
    const float* d_vectors;         /* optional [num_vectors * dim] FP32 rows */
    uint32_t num_vectors;
    uint32_t dim;
    uint32_t padded_dim;
} Fp16GpuTables;

#define FV_EMPTY_ID 0xFFFFFFFFu
#define FV_EMPTY_DIST 1e30f

#ifdef __cplusplus
}
static_assert(sizeof(GpuSearchResult) == 8 && alignof(GpuSearchResult) == 8, "GpuSearchResult layout");
static_assert(offsetof(GpuSearchResult, distance) == 4, "GpuSearchResult layout");
static_assert(sizeof(MetricType) == 4, "MetricType is a 32-bit C enum");
static_assert(sizeof(IvfPqGpuTables) == 72, "IvfPqGpuTables layout (LP64)");
static_assert(offsetof(IvfPqGpuTables, num_vectors) == 48, "IvfPqGpuTables layout");
static_assert(offsetof(IvfPqGpuTables, sub_dim) == 64, "IvfPqGpuTables layout");
static_assert(sizeof(Fp16GpuTables) == 40, "Fp16GpuTables layout (LP64)");
static_assert(offsetof(Fp16GpuTables, num_vectors) == 24, "Fp16GpuTables layout");
#else
_Static_assert(sizeof(GpuSearchResult) == 8, "GpuSearchResult layout");
_Static_assert(sizeof(IvfPqGpuTables) == 72, "IvfPqGpuTables layout (LP64)");
_Static_assert(sizeof(Fp16GpuTables) == 40, "Fp16GpuTables layout (LP64)");
#endif

#undef FV_ALIGN8
#endif /* FLASHVECTOR_AMD_TYPES_H */
