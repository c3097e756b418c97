// Autotuner of the matrix-core flat search (fv_flat_autotune).
//
// No counterpart in the reference (its launch shapes are fixed: one block
// per 16 queries, k <= 256, no shortlist). For one table, query sample, k,
// metric and dtype it chooses the shortlist depth that meets a recall target
// against the exact FP32 scan, then the work-item shape (row chunks per query
// block) and the first pass's row fraction by timing. A set-up call: it
// allocates its own device buffers and synchronizes the stream.
#include <algorithm>
#include <unordered_set>
#include <vector>

#include "../../include/flashvector/fv_ext.h"
#include "fv_flat_mfma.hpp"
#include "fv_ivfpq.hpp"

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

float recall_of(const std::vector<GpuSearchResult>& got, const std::vector<GpuSearchResult>& ex,
                uint32_t nq, uint32_t k) {
    size_t hits = 0, total = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        std::unordered_set<uint32_t> truth;
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t id = ex[(size_t)i * k + j].id;
            if (id != FV_EMPTY_ID) truth.insert(id);
        }
        total += truth.size();
        for (uint32_t j = 0; j < k; ++j) hits += truth.count(got[(size_t)i * k + j].id);
    }
    return total ? (float)hits / (float)total : 1.0f;
}

}  // namespace

extern "C" int fv_flat_autotune(const FvFlatTables* t, const float* d_queries, uint32_t nq,
                                const FvFlatSearchParams* base, float target_recall,
                                uint32_t trials, FvFlatTuning* out_best, void* stream) {
    if (!t || !d_queries || !base || !out_best || !t->d_vectors) return -1;
    if (nq == 0 || base->top_k == 0 || t->num_vectors == 0) return (int)hipErrorInvalidValue;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t k = base->top_k;
    const size_t res_bytes = (size_t)nq * k * sizeof(GpuSearchResult);
    DevBuf ws, res;
    hipError_t e = res.ensure(res_bytes);
    if (e != hipSuccess) return (int)e;
    // Exact FP32 top-k of the sample.
    const size_t exw = fv_flat_exact_workspace_size(t->num_vectors, t->dim, nq, k,
                                                    (MetricType)base->metric);
    if (exw == 0) return (int)hipErrorInvalidValue;
    if ((e = ws.ensure(exw)) != hipSuccess) return (int)e;
    int rc = fv_flat_exact_search(t->d_vectors, t->num_vectors, t->dim, d_queries, nq, k,
                                  (MetricType)base->metric, base->d_valid_mask,
                                  static_cast<GpuSearchResult*>(res.p), ws.p, ws.bytes, stream);
    if (rc) return rc;
    std::vector<GpuSearchResult> exact((size_t)nq * k), got((size_t)nq * k);
    if ((e = hipMemcpyAsync(exact.data(), res.p, res_bytes, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return (int)e;

    auto run = [&](const FvFlatSearchParams& p) -> int {
        const size_t need = fv_flat_workspace_size(t, nq, &p);
        if (need == 0) return (int)hipErrorInvalidValue;
        hipError_t ee = ws.ensure(need);
        if (ee != hipSuccess) return (int)ee;
        return fv_flat_search(t, d_queries, nq, &p, static_cast<GpuSearchResult*>(res.p), ws.p,
                              ws.bytes, stream);
    };

    // (1) Shortlist depth: the smallest factor that meets the recall target,
    // unless the caller fixed it (then only its recall is measured).
    FvFlatSearchParams p = *base;
    p.shortlist = 0;
    p.n_chunks = 0;
    p.bound_div = 0;
    p.d_top_k = nullptr;
    float recall = 0.0f;
    uint32_t factor = 16;
    const bool fixed = base->shortlist_factor != 0;
    const std::vector<uint32_t> factors =
        fixed ? std::vector<uint32_t>{base->shortlist_factor} : std::vector<uint32_t>{1u, 2u, 4u, 8u, 16u};
    for (uint32_t f : factors) {
        p.shortlist_factor = f;
        if ((rc = run(p)) != 0) return rc;
        if ((e = hipMemcpyAsync(got.data(), res.p, res_bytes, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return (int)e;
        recall = recall_of(got, exact, nq, k);
        factor = f;
        if (recall >= target_recall) break;
    }
    p.shortlist_factor = factor;

    // (2) Work shape and first-pass fraction, by time.
    fv::FlatMfmaPlan ap;
    if (!fv::flat_plan_for(t, nq, &p, &ap)) return (int)hipErrorInvalidValue;
    const uint32_t rows_plan = ap.compact && !ap.dyn_rows ? base->valid_count : t->num_vectors;
    const uint32_t auto_div = ap.bound_rows ? std::max(1u, rows_plan / ap.bound_rows) : 1u;
    std::vector<uint32_t> chunks = {std::max(1u, ap.n_chunks / 2), ap.n_chunks, ap.n_chunks * 2,
                                    ap.n_chunks * 4, ap.n_chunks * 8};
    std::sort(chunks.begin(), chunks.end());
    chunks.erase(std::unique(chunks.begin(), chunks.end()), chunks.end());
    std::vector<uint32_t> divs = {0u};
    if (ap.bound_chunks) divs = {0u, auto_div * 2, auto_div * 4, 0xFFFFFFFFu};
    hipEvent_t e0, e1;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return (int)e;
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return (int)e;
    }
    // Each candidate: median of `trials` timed runs after a warm-up. The
    // automatic plan (n_chunks 0, bound_div 0) is kept unless a candidate is
    // at least 3% faster: smaller gaps are within run-to-run noise.
    auto time_plan = [&](uint32_t c, uint32_t dv, float* out_ms) -> int {
        p.n_chunks = c;
        p.bound_div = dv;
        int r = run(p);  // warm-up
        std::vector<float> ts;
        for (uint32_t t = 0; t < std::max(1u, trials) && r == 0; ++t) {
            (void)hipEventRecord(e0, s);
            r = run(p);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ts.push_back(ms);
        }
        if (r == 0) {
            std::sort(ts.begin(), ts.end());
            *out_ms = ts[ts.size() / 2];
        }
        return r;
    };
    FvFlatTuning best{};
    float auto_ms = 0.0f;
    rc = time_plan(0u, 0u, &auto_ms);
    best.ms = auto_ms;
    for (uint32_t c : chunks) {
        for (uint32_t dv : divs) {
            float ms = 0.0f;
            if (rc == 0) rc = time_plan(c, dv, &ms);
            if (rc) break;
            if (ms < best.ms && ms < 0.97f * auto_ms) {
                best.ms = ms;
                best.n_chunks = c;
                best.bound_div = dv;
            }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    best.shortlist_factor = factor;
    best.recall = recall;
    *out_best = best;
    return 0;
}

// Autotuner of the IVF-PQ work-item plan (fv_ivf_pq_autotune). Every plan
// returns the same results, so only time decides.
extern "C" int fv_ivf_pq_autotune(const IvfPqGpuTables* t, const uint8_t* d_codes_list_order,
                                  const float* d_queries, uint32_t nq, const uint32_t* d_top_k,
                                  uint32_t max_top_k, const uint32_t* d_nprobe,
                                  uint32_t max_nprobe, uint32_t shortlist, int rerank,
                                  MetricType metric, const uint32_t* d_valid_mask,
                                  uint32_t trials, FvIvfPqTuning* out_best, void* stream) {
    if (!t || !d_queries || !out_best) return -1;
    if (nq == 0) return (int)hipErrorInvalidValue;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t need = fv_ivf_pq_workspace_size(t, nq, max_nprobe, shortlist, max_top_k);
    if (need == 0) return (int)hipErrorInvalidValue;
    DevBuf ws, res;
    hipError_t e = ws.ensure(need);
    if (e == hipSuccess) e = res.ensure((size_t)nq * max_top_k * sizeof(GpuSearchResult));
    if (e != hipSuccess) return (int)e;
    hipEvent_t e0, e1;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return (int)e;
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return (int)e;
    }
    auto run = [&](const FvIvfPqTuning* tu) {
        return fv_ivf_pq_search_tuned(t, d_codes_list_order, d_queries, nq, d_top_k, max_top_k,
                                      d_nprobe, max_nprobe, shortlist, rerank, metric,
                                      d_valid_mask, static_cast<GpuSearchResult*>(res.p), ws.p,
                                      ws.bytes, tu, stream);
    };
    auto time_plan = [&](const FvIvfPqTuning* tu, float* out_ms) -> int {
        int r = run(tu);  // warm-up
        std::vector<float> ts;
        for (uint32_t i = 0; i < std::max(1u, trials) && r == 0; ++i) {
            (void)hipEventRecord(e0, s);
            r = run(tu);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ts.push_back(ms);
        }
        if (r == 0) {
            std::sort(ts.begin(), ts.end());
            *out_ms = ts[ts.size() / 2];
        }
        return r;
    };
    float auto_ms = 0.0f;
    int rc = time_plan(nullptr, &auto_ms);
    FvIvfPqTuning best{};
    best.target_items = fv::kDefaultTargetItems;
    best.chunk_min = fv::kDefaultChunkMin;
    best.ms = auto_ms;
    for (uint32_t items : {0u, 256u, 1024u, 4096u, 16384u}) {
        for (uint32_t cmin : {4096u, 16384u, 65536u}) {
            if (rc) break;
            if (items == 0u && cmin != 16384u) continue;  // chunk_min is unused without splitting
            FvIvfPqTuning c{};
            c.target_items = items;
            c.chunk_min = cmin;
            float ms = 0.0f;
            rc = time_plan(&c, &ms);
            if (rc == 0 && ms < best.ms && ms < 0.97f * auto_ms) {
                best = c;
                best.ms = ms;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *out_best = best;
    return 0;
}
