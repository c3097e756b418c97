/*
 * CPU oracle for the FlashVector query-scan hot path (test infrastructure).
 * Build flags matter: -O2 -ffp-contract=off, no -ffast-math (see Makefile), so
 * every a*b+c below rounds twice unless written as fmaf().
 */
#include "fv_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LEGACY_MAX_TOPK 256u
#define LEGACY_MAX_SHORTLIST 10240u
#define EXT_MAX_TOPK 1024u
#define EXT_MAX_SHORTLIST 16384u

/* ------------------------------------------------------------------------ */
/* (distance, id) bounded selection                                          */
/* ------------------------------------------------------------------------ */

static inline int res_less(GpuSearchResult a, GpuSearchResult b) {
    return a.distance < b.distance || (a.distance == b.distance && a.id < b.id);
}

/* Max-heap of the `cap` smallest entries seen so far. */
typedef struct {
    GpuSearchResult* h;
    uint32_t size, cap;
} TopHeap;

static void heap_push(TopHeap* t, GpuSearchResult r) {
    if (!(r.distance < FV_EMPTY_DIST)) return; /* sentinel admission rule */
    if (t->cap == 0) return;
    if (t->size < t->cap) {
        uint32_t i = t->size++;
        while (i > 0) {
            uint32_t p = (i - 1) / 2;
            if (!res_less(t->h[p], r)) break;
            t->h[i] = t->h[p];
            i = p;
        }
        t->h[i] = r;
        return;
    }
    if (!res_less(r, t->h[0])) return;
    uint32_t i = 0;
    for (;;) {
        uint32_t l = 2 * i + 1, rr = l + 1, m = i;
        GpuSearchResult best = r;
        if (l < t->size && res_less(best, t->h[l])) { m = l; best = t->h[l]; }
        if (rr < t->size && res_less(best, t->h[rr])) { m = rr; }
        if (m == i) break;
        t->h[i] = t->h[m];
        i = m;
    }
    t->h[i] = r;
}

static int cmp_res(const void* a, const void* b) {
    GpuSearchResult x = *(const GpuSearchResult*)a, y = *(const GpuSearchResult*)b;
    if (res_less(x, y)) return -1;
    if (res_less(y, x)) return 1;
    return 0;
}

/* Sorts the heap contents ascending into out[0..size) and pads to `width`. */
static void heap_emit(TopHeap* t, GpuSearchResult* out, uint32_t width) {
    qsort(t->h, t->size, sizeof(GpuSearchResult), cmp_res);
    for (uint32_t i = 0; i < width; ++i) {
        if (i < t->size) out[i] = t->h[i];
        else { out[i].id = FV_EMPTY_ID; out[i].distance = FV_EMPTY_DIST; }
    }
}

static inline int mask_allows(const uint32_t* mask, uint32_t idx) {
    /* kernels/src/distance_metrics.cuh:17-20: LSB-first u32 words, NULL = all */
    return mask == NULL || ((mask[idx >> 5] >> (idx & 31u)) & 1u) != 0u;
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* ------------------------------------------------------------------------ */
/* Distances                                                                 */
/* ------------------------------------------------------------------------ */

float fvo_distance(const float* q, const float* x, uint32_t dim, int metric) {
    if (metric == METRIC_INNER_PRODUCT) {
        float dot = 0.0f;
        for (uint32_t d = 0; d < dim; ++d) {
            float p = q[d] * x[d];
            dot = dot + p;
        }
        return -dot;
    }
    if (metric == METRIC_COSINE) {
        float dot = 0.0f, nq = 0.0f, nx = 0.0f;
        for (uint32_t d = 0; d < dim; ++d) {
            float a = q[d], b = x[d];
            float p = a * b, pa = a * a, pb = b * b;
            dot = dot + p;
            nq = nq + pa;
            nx = nx + pb;
        }
        float denom = sqrtf(nq) * sqrtf(nx) + 1e-8f;
        return 1.0f - dot / denom;
    }
    float acc = 0.0f;
    for (uint32_t d = 0; d < dim; ++d) {
        float diff = q[d] - x[d];
        float sq = diff * diff;
        acc = acc + sq;
    }
    return acc;
}

void fvo_distance_matrix(const float* queries, const float* dataset, uint32_t nq,
                         uint32_t n, uint32_t dim, int metric, float* out) {
    for (uint32_t i = 0; i < nq; ++i)
        for (uint32_t j = 0; j < n; ++j)
            out[(size_t)i * n + j] =
                fvo_distance(queries + (size_t)i * dim, dataset + (size_t)j * dim, dim, metric);
}

uint32_t fvo_exact_knn(const float* dataset, uint32_t n, uint32_t dim, const float* query,
                       uint32_t top_k, GpuSearchResult* out) {
    /* metrics.rs keeps every distance (no sentinel), so no admission rule here. */
    GpuSearchResult* all = (GpuSearchResult*)malloc(sizeof(GpuSearchResult) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) {
        all[i].id = i;
        all[i].distance = fvo_distance(query, dataset + (size_t)i * dim, dim, METRIC_L2);
    }
    qsort(all, n, sizeof(GpuSearchResult), cmp_res);
    uint32_t cnt = top_k < n ? top_k : n;
    memcpy(out, all, sizeof(GpuSearchResult) * cnt);
    free(all);
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* Flat scans                                                                */
/* ------------------------------------------------------------------------ */

static void flat_one_query(const float* vectors, const uint32_t* vec_ids, uint32_t n,
                           uint32_t dim, const float* q, uint32_t top_k, int metric,
                           const uint32_t* mask, GpuSearchResult* out, GpuSearchResult* heapbuf) {
    TopHeap t = {heapbuf, 0, top_k};
    for (uint32_t s = 0; s < n; ++s) {
        if (!mask_allows(mask, s)) continue;
        GpuSearchResult r;
        r.id = s; /* ties by slot; remapped after selection */
        r.distance = fvo_distance(q, vectors + (size_t)s * dim, dim, metric);
        heap_push(&t, r);
    }
    heap_emit(&t, out, top_k);
    if (vec_ids) {
        for (uint32_t i = 0; i < t.size; ++i) out[i].id = vec_ids[out[i].id];
    }
}

int fvo_flat_scan_batch(const float* vectors, const uint32_t* vec_ids, uint32_t n,
                        uint32_t dim, const float* queries, uint32_t nq, uint32_t top_k,
                        int metric, const uint32_t* valid_mask, GpuSearchResult* out,
                        int nthreads) {
    if (!vectors || !queries || !out) return -1; /* sparse_hybrid.cu:517-518 */
    if (nq == 0 || top_k == 0) return 0;
    if (top_k > LEGACY_MAX_TOPK) top_k = LEGACY_MAX_TOPK;
    if (n == 0) {
        memset(out, 0xFF, sizeof(GpuSearchResult) * (size_t)nq * top_k);
        return 0;
    }
    set_threads(nthreads);
#pragma omp parallel
    {
        GpuSearchResult* hb = (GpuSearchResult*)malloc(sizeof(GpuSearchResult) * top_k);
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < (int64_t)nq; ++i)
            flat_one_query(vectors, vec_ids, n, dim, queries + (size_t)i * dim, top_k, metric,
                           valid_mask, out + (size_t)i * top_k, hb);
        free(hb);
    }
    return 0;
}

int fvo_flat_topk(const float* vectors, uint32_t n, uint32_t dim, const float* queries,
                  uint32_t nq, uint32_t top_k, int metric, const uint32_t* valid_mask,
                  GpuSearchResult* out, int nthreads) {
    if (!vectors || !queries || !out) return -1;
    if (nq == 0 || top_k == 0) return 0;
    set_threads(nthreads);
#pragma omp parallel
    {
        GpuSearchResult* hb = (GpuSearchResult*)malloc(sizeof(GpuSearchResult) * top_k);
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < (int64_t)nq; ++i)
            flat_one_query(vectors, NULL, n, dim, queries + (size_t)i * dim, top_k, metric,
                           valid_mask, out + (size_t)i * top_k, hb);
        free(hb);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* IVF-PQ                                                                    */
/* ------------------------------------------------------------------------ */

void fvo_coarse_probe(const float* query, const float* centroids, uint32_t dim,
                      uint32_t nlist, uint32_t nprobe, uint32_t* probed, float* probed_dist) {
    GpuSearchResult* hb = (GpuSearchResult*)malloc(sizeof(GpuSearchResult) * (nprobe ? nprobe : 1));
    TopHeap t = {hb, 0, nprobe};
    for (uint32_t c = 0; c < nlist; ++c) {
        const float* cen = centroids + (size_t)c * dim;
        float dist = 0.0f;
        for (uint32_t d = 0; d < dim; ++d) {
            float diff = query[d] - cen[d];
            dist = fmaf(diff, diff, dist); /* ivf_rerank.cuh:81-84 */
        }
        GpuSearchResult r = {c, dist};
        heap_push(&t, r);
    }
    qsort(t.h, t.size, sizeof(GpuSearchResult), cmp_res);
    for (uint32_t p = 0; p < nprobe; ++p) {
        probed[p] = p < t.size ? t.h[p].id : FV_EMPTY_ID;
        if (probed_dist) probed_dist[p] = p < t.size ? t.h[p].distance : FV_EMPTY_DIST;
    }
    free(hb);
}

void fvo_pq_lut(const float* query, const float* centroid, const float* codebooks,
                uint32_t m_pq, uint32_t sub_dim, float* lut) {
    for (uint32_t m = 0; m < m_pq; ++m) {
        for (uint32_t j = 0; j < 256u; ++j) {
            const float* cb = codebooks + ((size_t)m * 256u + j) * sub_dim;
            float dist = 0.0f;
            for (uint32_t s = 0; s < sub_dim; ++s) {
                float r = query[m * sub_dim + s] - centroid[m * sub_dim + s];
                float diff = r - cb[s];
                dist = fmaf(diff, diff, dist); /* ivf_pq_lookup.cu:157-158 */
            }
            lut[(size_t)m * 256u + j] = dist;
        }
    }
}

/* One query: coarse probe, LUT, ADC shortlist of depth S, optional rerank,
 * then the first k by (distance, id) into out[0..width). */
static void ivfpq_one_query(const IvfPqGpuTables* t, const float* q, uint32_t k,
                            uint32_t nprobe, uint32_t S, int do_rerank, int metric,
                            const uint32_t* mask, GpuSearchResult* out, uint32_t width,
                            float* lut, uint32_t* probed, GpuSearchResult* heapbuf) {
    fvo_coarse_probe(q, t->d_centroids, t->dim, t->nlist, nprobe, probed, NULL);
    TopHeap h = {heapbuf, 0, S};
    for (uint32_t p = 0; p < nprobe; ++p) {
        uint32_t c = probed[p];
        if (c >= t->nlist) continue;
        fvo_pq_lut(q, t->d_centroids + (size_t)c * t->dim, t->d_pq_codebooks, t->m_pq,
                   t->sub_dim, lut);
        uint32_t beg = t->d_ivf_offsets[c], end = t->d_ivf_offsets[c + 1];
        for (uint32_t i = beg; i < end; ++i) {
            uint32_t id = t->d_ivf_vec_ids[i];
            if (!mask_allows(mask, id)) continue;
            const uint8_t* code = t->d_pq_codes + (size_t)id * t->m_pq;
            float adc = 0.0f;
            for (uint32_t m = 0; m < t->m_pq; ++m) adc = adc + lut[(size_t)m * 256u + code[m]];
            GpuSearchResult r = {id, adc};
            heap_push(&h, r);
        }
    }
    if (do_rerank) {
        for (uint32_t i = 0; i < h.size; ++i) {
            uint32_t id = h.h[i].id;
            h.h[i].distance = id < t->num_vectors
                ? fvo_distance(q, t->d_vectors + (size_t)id * t->dim, t->dim, metric)
                : FV_EMPTY_DIST;
        }
    }
    qsort(h.h, h.size, sizeof(GpuSearchResult), cmp_res);
    for (uint32_t i = 0; i < width; ++i) {
        if (i < k && i < h.size) out[i] = h.h[i];
        else { out[i].id = FV_EMPTY_ID; out[i].distance = FV_EMPTY_DIST; }
    }
}

int fvo_ivf_pq_search_batch(const IvfPqGpuTables* t, const float* queries, uint32_t nq,
                            uint32_t top_k, uint32_t nprobe, uint32_t rerank_k, int metric,
                            const uint32_t* valid_mask, GpuSearchResult* out, int nthreads) {
    if (!t || !queries || !out) return -1; /* ivf_pq_lookup.cu:462-464 */
    if (nq == 0 || top_k == 0 || t->m_pq == 0 || t->nlist == 0) return 0;
    if (top_k > LEGACY_MAX_TOPK) top_k = LEGACY_MAX_TOPK;
    uint32_t np = nprobe == 0 ? 1u : nprobe;
    if (np > t->nlist) np = t->nlist;
    uint32_t S = rerank_k == 0 ? top_k : rerank_k;
    if (S < top_k) S = top_k;
    if (S > LEGACY_MAX_SHORTLIST) S = LEGACY_MAX_SHORTLIST;
    int do_rerank = rerank_k != 1u && t->d_vectors != NULL;
    set_threads(nthreads);
#pragma omp parallel
    {
        float* lut = (float*)malloc(sizeof(float) * (size_t)t->m_pq * 256u);
        uint32_t* probed = (uint32_t*)malloc(sizeof(uint32_t) * np);
        GpuSearchResult* hb = (GpuSearchResult*)malloc(sizeof(GpuSearchResult) * S);
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < (int64_t)nq; ++i)
            ivfpq_one_query(t, queries + (size_t)i * t->dim, top_k, np, S, do_rerank, metric,
                            valid_mask, out + (size_t)i * top_k, top_k, lut, probed, hb);
        free(lut);
        free(probed);
        free(hb);
    }
    return 0;
}

int fvo_ivf_pq_search_uneven(const IvfPqGpuTables* t, const float* queries, uint32_t nq,
                             const uint32_t* k_per_query, const uint32_t* nprobe_per_query,
                             uint32_t shortlist, int rerank, int metric,
                             const uint32_t* valid_mask, GpuSearchResult* out,
                             uint32_t out_stride, int nthreads) {
    if (!t || !queries || !out || !k_per_query || !nprobe_per_query) return -1;
    if (nq == 0 || t->m_pq == 0 || t->nlist == 0) return 0;
    int do_rerank = rerank && t->d_vectors != NULL;
    set_threads(nthreads);
#pragma omp parallel
    {
        float* lut = (float*)malloc(sizeof(float) * (size_t)t->m_pq * 256u);
        uint32_t* probed = (uint32_t*)malloc(sizeof(uint32_t) * t->nlist);
        GpuSearchResult* hb = (GpuSearchResult*)malloc(sizeof(GpuSearchResult) * EXT_MAX_SHORTLIST);
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < (int64_t)nq; ++i) {
            uint32_t k = k_per_query[i];
            if (k > EXT_MAX_TOPK) k = EXT_MAX_TOPK;
            if (k > out_stride) k = out_stride;
            uint32_t np = nprobe_per_query[i] == 0 ? 1u : nprobe_per_query[i];
            if (np > t->nlist) np = t->nlist;
            uint32_t S = shortlist < k ? k : shortlist;
            if (S > EXT_MAX_SHORTLIST) S = EXT_MAX_SHORTLIST;
            ivfpq_one_query(t, queries + (size_t)i * t->dim, k, np, S, do_rerank, metric,
                            valid_mask, out + (size_t)i * out_stride, out_stride, lut, probed, hb);
        }
        free(lut);
        free(probed);
        free(hb);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Merge                                                                     */
/* ------------------------------------------------------------------------ */

int fvo_merge_topk(const GpuSearchResult* shard_results, uint32_t num_shards, uint32_t nq,
                   uint32_t shard_top_k, uint32_t out_top_k, GpuSearchResult* out) {
    if (!shard_results || !out) return -1; /* sparse_hybrid.cu:536-537 */
    if (num_shards == 0 || nq == 0 || out_top_k == 0) return 0;
    if (out_top_k > LEGACY_MAX_TOPK) out_top_k = LEGACY_MAX_TOPK;
    uint64_t total = (uint64_t)num_shards * shard_top_k;
    if (total > 4096u) total = 4096u; /* sparse_hybrid.cu:211-212 buffer cap */
    GpuSearchResult* buf = (GpuSearchResult*)malloc(sizeof(GpuSearchResult) * (total ? total : 1));
    for (uint32_t q = 0; q < nq; ++q) {
        uint32_t cnt = 0;
        for (uint32_t i = 0; i < (uint32_t)total; ++i) {
            uint32_t s = i / shard_top_k, kk = i % shard_top_k;
            GpuSearchResult r = shard_results[((size_t)s * nq + q) * shard_top_k + kk];
            if (r.id != FV_EMPTY_ID) buf[cnt++] = r;
        }
        qsort(buf, cnt, sizeof(GpuSearchResult), cmp_res);
        GpuSearchResult* o = out + (size_t)q * out_top_k;
        uint32_t written = 0;
        for (uint32_t i = 0; i < cnt && written < out_top_k; ++i) {
            int dup = 0;
            for (uint32_t j = 0; j < written; ++j)
                if (o[j].id == buf[i].id) { dup = 1; break; }
            if (!dup) o[written++] = buf[i];
        }
        for (; written < out_top_k; ++written) {
            o[written].id = FV_EMPTY_ID;
            o[written].distance = FV_EMPTY_DIST;
        }
    }
    free(buf);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* binary16                                                                  */
/* ------------------------------------------------------------------------ */

uint16_t fvo_f32_to_f16_bits(float x) {
    uint32_t bits;
    memcpy(&bits, &x, 4);
    uint16_t sign = (uint16_t)((bits >> 16) & 0x8000u);
    int32_t exp = (int32_t)((bits >> 23) & 0xffu);
    uint32_t mant = bits & 0x007fffffu;
    if (exp == 255) return (uint16_t)(sign | (mant == 0 ? 0x7c00u : 0x7e00u));
    int32_t e16 = exp - 127 + 15;
    if (e16 >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e16 <= 0) {
        if (e16 < -10) return sign;
        uint32_t m = (mant | 0x00800000u) >> (14 - e16); /* truncating subnormal */
        return (uint16_t)(sign | (uint16_t)m);
    }
    uint32_t rounded = mant + 0x1000u; /* half-up on magnitude */
    if (rounded & 0x00800000u) {
        int32_t ne = e16 + 1;
        if (ne >= 31) return (uint16_t)(sign | 0x7c00u);
        return (uint16_t)(sign | (uint16_t)(ne << 10));
    }
    return (uint16_t)(sign | (uint16_t)(e16 << 10) | (uint16_t)(rounded >> 13));
}

float fvo_f16_bits_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1fu;
    uint32_t mant = h & 0x03ffu;
    uint32_t bits;
    if (exp == 0) {
        if (mant == 0) {
            bits = sign;
        } else {
            int32_t e = -14;
            uint32_t m = mant;
            while ((m & 0x0400u) == 0) { m <<= 1; e -= 1; }
            m &= 0x03ffu;
            bits = sign | ((uint32_t)(e + 127) << 23) | (m << 13);
        }
    } else if (exp == 31) {
        bits = sign | (0xffu << 23) | (mant << 13);
    } else {
        bits = sign | ((exp + 112u) << 23) | (mant << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

uint32_t fvo_encode_fp16_padded(const float* vectors, uint32_t n, uint32_t dim,
                                uint16_t* out_fp16, float* out_norms) {
    uint32_t pd = (dim + 15u) / 16u * 16u;
    for (uint32_t i = 0; i < n; ++i) {
        const float* src = vectors + (size_t)i * dim;
        uint16_t* dst = out_fp16 + (size_t)i * pd;
        float nsq = 0.0f;
        for (uint32_t d = 0; d < pd; ++d) {
            if (d < dim) {
                float v = src[d];
                float sq = v * v;
                nsq = nsq + sq;
                dst[d] = fvo_f32_to_f16_bits(v);
            } else {
                dst[d] = 0;
            }
        }
        out_norms[i] = nsq;
    }
    return pd;
}

/* ------------------------------------------------------------------------ */
/* FP8 e4m3 and the FP64 top-k                                               */
/* ------------------------------------------------------------------------ */

uint8_t fvo_f32_to_e4m3(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    const uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    u &= 0x7fffffffu;
    if (u > 0x7f800000u) return (uint8_t)(sign | 0x7fu);
    float a;
    memcpy(&a, &u, 4);
    if (a >= 448.0f) return (uint8_t)(sign | 0x7eu);
    const int e = (int)(u >> 23) - 127;
    if (e < -6) return (uint8_t)(sign | (uint32_t)rintf(a * 512.0f)); /* step 2^-9 */
    const uint32_t mant = u & 0x7fffffu;
    uint32_t r = mant >> 20;
    const uint32_t rem = mant & 0xfffffu;
    if (rem > 0x80000u || (rem == 0x80000u && (r & 1u))) ++r;
    int ef = e + 7;
    if (r == 8u) {
        r = 0u;
        ++ef;
    }
    if (ef > 15 || (ef == 15 && r == 7u)) return (uint8_t)(sign | 0x7eu);
    return (uint8_t)(sign | (uint32_t)(ef << 3) | r);
}

float fvo_e4m3_to_f32(uint8_t c) {
    const uint32_t e = (c >> 3) & 15u, m = c & 7u;
    float v;
    if (e == 15u && m == 7u) v = NAN;
    else if (e == 0u) v = (float)m / 512.0f;
    else v = ldexpf(1.0f + (float)m / 8.0f, (int)e - 7);
    return (c & 0x80u) ? -v : v;
}

void fvo_encode_fp8_rows(const float* x, uint32_t n, uint32_t dim, uint32_t pdim, float sign,
                         uint8_t* codes, float* scales) {
    for (uint32_t r = 0; r < n; ++r) {
        const float* row = x + (size_t)r * dim;
        float m = 0.0f;
        for (uint32_t d = 0; d < dim; ++d) m = fmaxf(m, fabsf(row[d]));
        const float s = m > 0.0f ? m / 448.0f : 1.0f;
        for (uint32_t d = 0; d < pdim; ++d)
            codes[(size_t)r * pdim + d] = d < dim ? fvo_f32_to_e4m3(sign * row[d] / s) : 0u;
        if (scales) scales[r] = s;
    }
}

typedef struct {
    double d;
    uint32_t id;
} D64Res;

static int cmp_d64(const void* a, const void* b) {
    const D64Res* x = (const D64Res*)a;
    const D64Res* y = (const D64Res*)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

static double dist_f64(const float* q, const float* x, uint32_t dim, int metric) {
    double a = 0.0, nq = 0.0, nx = 0.0;
    for (uint32_t d = 0; d < dim; ++d) {
        const double qd = q[d], xd = x[d];
        if (metric == METRIC_L2) {
            const double t = qd - xd;
            a = fma(t, t, a);
        } else {
            a = fma(qd, xd, a);
            nq = fma(qd, qd, nq);
            nx = fma(xd, xd, nx);
        }
    }
    if (metric == METRIC_L2) return a;
    if (metric == METRIC_INNER_PRODUCT) return -a;
    return 1.0 - a / (sqrt(nq) * sqrt(nx) + 1e-8);
}

int fvo_flat_topk_f64(const float* vectors, uint32_t n, uint32_t dim, const float* queries,
                      uint32_t nq, uint32_t top_k, int metric, const uint32_t* valid_mask,
                      uint32_t* out_ids, double* out_dist, int nthreads) {
    if (!vectors || !queries || !out_ids || !out_dist) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (uint32_t i = 0; i < nq; ++i) {
        D64Res* all = (D64Res*)malloc(sizeof(D64Res) * (n ? n : 1));
        uint32_t c = 0;
        for (uint32_t r = 0; r < n; ++r) {
            if (valid_mask && !((valid_mask[r >> 5] >> (r & 31u)) & 1u)) continue;
            all[c].d = dist_f64(queries + (size_t)i * dim, vectors + (size_t)r * dim, dim, metric);
            all[c].id = r;
            ++c;
        }
        qsort(all, c, sizeof(D64Res), cmp_d64);
        for (uint32_t j = 0; j < top_k; ++j) {
            out_ids[(size_t)i * top_k + j] = j < c ? all[j].id : 0xFFFFFFFFu;
            out_dist[(size_t)i * top_k + j] = j < c ? all[j].d : INFINITY;
        }
        free(all);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* IVF-PQ index construction                                                 */
/* ------------------------------------------------------------------------ */

/* splitmix64 (Steele, Lea, Flood 2014; Vigna's public-domain constants). */
static uint64_t sm64_next(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint32_t sm64_below(uint64_t* s, uint32_t m) { return (uint32_t)(sm64_next(s) % m); }
static double sm64_unit(uint64_t* s) { return (double)(sm64_next(s) >> 11) * (1.0 / 9007199254740992.0); }

/* KMeans::l2_sq (cluster.rs:18-26): sequential, non-fused. */
static float km_l2(const float* a, const float* b, uint32_t dim) {
    float acc = 0.0f;
    for (uint32_t d = 0; d < dim; ++d) {
        float t = a[d] - b[d];
        float sq = t * t;
        acc = acc + sq;
    }
    return acc;
}

/* Assign step (cluster.rs:97-123): first strict minimum, best 0 if none. */
void fvo_kmeans_assign(const float* x, uint32_t n, uint32_t dim, uint32_t ld,
                       const float* centroids, uint32_t k, uint32_t* assign) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float* v = x + (size_t)i * ld;
        float best = INFINITY;
        uint32_t bc = 0;
        for (uint32_t c = 0; c < k; ++c) {
            float d = km_l2(v, centroids + (size_t)c * dim, dim);
            if (d < best) {
                best = d;
                bc = c;
            }
        }
        assign[i] = bc;
    }
}

#define KM_SEED_CHUNK 256u

/* k-means++ over a bounded sample (cluster.rs:46-93), deterministic. */
static void km_seed(const float* x, uint32_t n, uint32_t dim, uint32_t ld, uint32_t k,
                    uint64_t* rng, float* cent) {
    uint32_t pool = k * 64u < n ? k * 64u : n;
    if (pool == 0) pool = 1;
    uint32_t* idx = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    for (uint32_t i = 0; i < pool; ++i) {
        uint32_t j = i + sm64_below(rng, n - i);
        uint32_t t = idx[i]; idx[i] = idx[j]; idx[j] = t;
    }
    double* u = (double*)malloc((size_t)(k > 1 ? k : 1) * sizeof(double));
    for (uint32_t c = 1; c < k; ++c) u[c] = sm64_unit(rng);
    float* closest = (float*)malloc((size_t)pool * sizeof(float));
    uint32_t nch = (pool + KM_SEED_CHUNK - 1) / KM_SEED_CHUNK;
    double* cs = (double*)malloc((size_t)nch * sizeof(double));
    for (uint32_t i = 0; i < pool; ++i) closest[i] = INFINITY;
    memcpy(cent, x + (size_t)idx[0] * ld, dim * sizeof(float));
    for (uint32_t c = 1; c < k; ++c) {
        const float* last = cent + (size_t)(c - 1) * dim;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)pool; ++i) {
            float cand = km_l2(x + (size_t)idx[i] * ld, last, dim);
            if (cand < closest[i]) closest[i] = cand;
        }
        double total = 0.0;
        for (uint32_t j = 0; j < nch; ++j) {
            double s = 0.0;
            uint32_t e = (j + 1) * KM_SEED_CHUNK < pool ? (j + 1) * KM_SEED_CHUNK : pool;
            for (uint32_t i = j * KM_SEED_CHUNK; i < e; ++i) s = s + (double)closest[i];
            cs[j] = s;
            total = total + s;
        }
        uint32_t chosen = idx[pool - 1];
        if (total > 0.0) {
            double target = u[c] * total, run = 0.0;
            for (uint32_t j = 0; j < nch; ++j) {
                uint32_t e = (j + 1) * KM_SEED_CHUNK < pool ? (j + 1) * KM_SEED_CHUNK : pool;
                if (run + cs[j] > target) {
                    double r = run;
                    chosen = idx[e - 1];
                    for (uint32_t i = j * KM_SEED_CHUNK; i < e; ++i) {
                        r = r + (double)closest[i];
                        if (r > target) {
                            chosen = idx[i];
                            break;
                        }
                    }
                    break;
                }
                run = run + cs[j];
            }
        }
        memcpy(cent + (size_t)c * dim, x + (size_t)chosen * ld, dim * sizeof(float));
    }
    free(cs);
    free(closest);
    free(u);
    free(idx);
}

void fvo_kmeans_fit(const float* x, uint32_t n, uint32_t dim, uint32_t ld, uint32_t k,
                    uint32_t max_iters, uint64_t seed, float* cent, uint32_t* assign,
                    uint32_t* sizes, uint32_t* iters_run) {
    uint64_t rng = seed;
    uint32_t it_done = 0;
    km_seed(x, n, dim, ld, k, &rng, cent);
    memset(assign, 0, (size_t)n * sizeof(uint32_t));
    memset(sizes, 0, (size_t)k * sizeof(uint32_t));
    float* mean = NULL;
    float* ncent = (float*)malloc((size_t)k * dim * sizeof(float));
    for (uint32_t it = 0; it < max_iters; ++it) {
        fvo_kmeans_assign(x, n, dim, ld, cent, k, assign);
        /* Update step (cluster.rs:125-171): sums in vector-index order. */
        memset(ncent, 0, (size_t)k * dim * sizeof(float));
        memset(sizes, 0, (size_t)k * sizeof(uint32_t));
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t c = assign[i];
            sizes[c]++;
            const float* v = x + (size_t)i * ld;
            float* s = ncent + (size_t)c * dim;
            for (uint32_t d = 0; d < dim; ++d) s[d] = s[d] + v[d];
        }
        float max_delta = 0.0f;
        for (uint32_t c = 0; c < k; ++c) {
            float* s = ncent + (size_t)c * dim;
            if (sizes[c] > 0) {
                float inv = 1.0f / (float)sizes[c];
                const float* old = cent + (size_t)c * dim;
                float delta = 0.0f;
                for (uint32_t d = 0; d < dim; ++d) {
                    s[d] = s[d] * inv;
                    float diff = s[d] - old[d];
                    float sq = diff * diff;
                    delta = delta + sq;
                }
                if (delta > max_delta) max_delta = delta;
            } else {
                /* Empty-cluster recovery (cluster.rs:151-168): the row furthest
                 * from the data mean among min(n, 500) drawn rows. */
                if (!mean) {
                    mean = (float*)calloc(dim, sizeof(float));
                    for (uint32_t i = 0; i < n; ++i) {
                        const float* v = x + (size_t)i * ld;
                        for (uint32_t d = 0; d < dim; ++d) mean[d] = mean[d] + v[d];
                    }
                    float inv_n = 1.0f / (float)n;
                    for (uint32_t d = 0; d < dim; ++d) mean[d] = mean[d] * inv_n;
                }
                uint32_t far = sm64_below(&rng, n);
                float max_d = 0.0f;
                uint32_t draws = n < 500u ? n : 500u;
                for (uint32_t t = 0; t < draws; ++t) {
                    uint32_t si = sm64_below(&rng, n);
                    float d = km_l2(x + (size_t)si * ld, mean, dim);
                    if (d > max_d) {
                        max_d = d;
                        far = si;
                    }
                }
                memcpy(s, x + (size_t)far * ld, dim * sizeof(float));
            }
        }
        memcpy(cent, ncent, (size_t)k * dim * sizeof(float));
        it_done = it + 1;
        if (max_delta < 1e-4f) break;
    }
    if (iters_run) *iters_run = it_done;
    free(ncent);
    free(mean);
}

/* GpuVectorIndex::build, IVF-PQ part (index.rs:405-458): coarse k-means over
 * all rows, residuals against the assigned centroid, one k-means per
 * sub-space on the residuals (ProductQuantizer::train, cluster.rs:197-244;
 * unused codewords FLT_MAX), encode (cluster.rs:246-286) and posting lists
 * ascending by vector id. Sub-quantizer m draws from seed + (m+1)*golden. */
int fvo_ivf_pq_build(const float* x, uint32_t n, uint32_t dim, uint32_t nlist, uint32_t m_pq,
                     uint32_t coarse_iters, uint32_t pq_iters, uint64_t seed, float* centroids,
                     float* codebooks, uint8_t* codes, uint32_t* offsets, uint32_t* vec_ids,
                     uint32_t* iters_run) {
    if (n == 0 || dim == 0 || m_pq == 0 || dim % m_pq) return -1;
    if (nlist > n) nlist = n;
    if (nlist == 0) nlist = 1;
    const uint32_t sub = dim / m_pq;
    uint32_t* assign = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    uint32_t* sizes = (uint32_t*)malloc((size_t)(nlist > 256 ? nlist : 256) * sizeof(uint32_t));
    fvo_kmeans_fit(x, n, dim, dim, nlist, coarse_iters, seed, centroids, assign, sizes,
                   iters_run ? &iters_run[0] : NULL);
    float* res = (float*)malloc((size_t)n * dim * sizeof(float));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float* c = centroids + (size_t)assign[i] * dim;
        for (uint32_t d = 0; d < dim; ++d) res[(size_t)i * dim + d] = x[(size_t)i * dim + d] - c[d];
    }
    const uint32_t kpq = n < 256u ? n : 256u;
    uint32_t* sub_assign = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    for (uint32_t m = 0; m < m_pq; ++m) {
        float* cb = codebooks + (size_t)m * 256 * sub;
        fvo_kmeans_fit(res + (size_t)m * sub, n, sub, dim, kpq, pq_iters,
                       seed + (uint64_t)(m + 1) * 0x9E3779B97F4A7C15ull, cb, sub_assign, sizes,
                       iters_run ? &iters_run[1 + m] : NULL);
        for (uint32_t c = kpq; c < 256; ++c)
            for (uint32_t d = 0; d < sub; ++d) cb[(size_t)c * sub + d] = FLT_MAX;
        fvo_kmeans_assign(res + (size_t)m * sub, n, sub, dim, cb, 256, sub_assign);
        for (uint32_t i = 0; i < n; ++i) codes[(size_t)i * m_pq + m] = (uint8_t)sub_assign[i];
    }
    memset(offsets, 0, (size_t)(nlist + 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) offsets[assign[i] + 1]++;
    for (uint32_t c = 0; c < nlist; ++c) offsets[c + 1] += offsets[c];
    uint32_t* fill = (uint32_t*)calloc(nlist, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) vec_ids[offsets[assign[i]] + fill[assign[i]]++] = i;
    free(fill);
    free(sub_assign);
    free(res);
    free(sizes);
    free(assign);
    return 0;
}
