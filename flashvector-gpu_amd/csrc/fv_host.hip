// Host search engine (fvh_*): stream pool with pre-allocated per-stream
// scratch, device-side mask build, main + delta search on two streams,
// device merge. Counterpart of crates/engine/src/index.rs:1264-1487,
// 1979-2237 and crates/engine/src/streams.rs:22-143.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/flashvector/cuda_bridge.h"
#include "../../include/flashvector/fv_build.h"
#include "../../include/flashvector/fv_host.h"
#include "fv_flat_exact.hpp"

namespace fvh {

using fv::kWave;

// Slot-indexed validity words: bit s is set when slot s is alive (valid
// words) and, with a caller mask, its logical id's bit is set there
// (index.rs:1264-1316, evaluated per slot on the device). One thread per
// slot; each wavefront writes the two words of its 64 slots.
__global__ __launch_bounds__(256) void slot_mask_kernel(const uint32_t* __restrict__ valid,
                                                        const uint32_t* __restrict__ slot2lid,
                                                        uint32_t n,
                                                        const uint32_t* __restrict__ user,
                                                        uint32_t user_words,
                                                        uint32_t* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = s < n && ((valid[s >> 5] >> (s & 31u)) & 1u);
    if (ok && user) {
        const uint32_t lid = slot2lid ? slot2lid[s] : s;
        const uint32_t w = lid >> 5;
        ok = w < user_words && ((user[w] >> (lid & 31u)) & 1u);
    }
    const uint64_t b = __ballot(ok);
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t base = (s - lane) >> 5;  // first word of this wavefront
    const uint32_t nwords = (n + 31) / 32;
    if (lane == 0 && base < nwords) out[base] = (uint32_t)b;
    if (lane == 32 && base + 1 < nwords) out[base + 1] = (uint32_t)(b >> 32);
}

// Result slot -> logical id.
__global__ void remap_ids_kernel(GpuSearchResult* __restrict__ r, size_t count,
                                 const uint32_t* __restrict__ map, uint32_t map_n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t id = r[i].id;
    if (id != FV_EMPTY_ID && id < map_n) r[i].id = map[id];
}

template <typename T>
struct DevArray {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e == hipSuccess) n = count;
        else p = nullptr;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevArray() { release(); }
};

template <typename T>
struct PinnedArray {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocDefault);
        if (e == hipSuccess) n = count;
        else p = nullptr;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    ~PinnedArray() { release(); }
};

constexpr int kFlat = 1, kIvf = 2;

// Per-stream scratch (streams.rs:22-69): everything one search touches.
struct Scratch {
    hipStream_t stream = nullptr;
    hipEvent_t queries_ready = nullptr, delta_done = nullptr, free_ev = nullptr;
    bool used = false;
    std::mutex mu;
    DevArray<float> q;
    DevArray<GpuSearchResult> res, merge;
    DevArray<uint32_t> mask, delta_mask, user_mask;
    DevArray<uint8_t> ws, dws;
    PinnedArray<float> hq;
    PinnedArray<GpuSearchResult> hres;
    PinnedArray<uint32_t> hmask;
    ~Scratch() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (queries_ready) (void)hipEventDestroy(queries_ready);
        if (delta_done) (void)hipEventDestroy(delta_done);
        if (free_ev) (void)hipEventDestroy(free_ev);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

}  // namespace fvh

struct FvEngine {
    FvEngineConfig cfg{};
    std::vector<std::unique_ptr<fvh::Scratch>> pool;
    std::atomic<uint64_t> rr{0};
    hipStream_t delta_stream = nullptr;
    int mode = 0;  // bits: kFlat (flat matrix-core tables), kIvf (IVF-PQ tables)
    uint32_t n_main = 0;
    // flat tables
    fvh::DevArray<float> rows, norms, scales, enc_err;
    fvh::DevArray<double> norms64;
    fvh::DevArray<uint8_t> codes;
    FvFlatTables flat{};
    // IVF-PQ tables
    fvh::DevArray<float> centroids, codebooks;
    fvh::DevArray<uint8_t> pq_codes, pq_codes_lo;
    fvh::DevArray<uint32_t> offsets, vec_ids;
    IvfPqGpuTables ivf{};
    // ids, tombstones
    fvh::DevArray<uint32_t> slot2lid, valid;
    std::vector<uint32_t> h_slot2lid, h_valid;
    std::unordered_map<uint32_t, uint32_t> lid2slot;
    uint32_t main_dead = 0;
    uint32_t max_lid = 0;
    // Deleted logical ids (index.rs:146 tombstones) and the next id that
    // fvh_add_vectors hands out (index.rs:145 next_logical_id). Logical ids are
    // unique over the main index, the delta buffer and the tombstones.
    std::unordered_set<uint32_t> tombstones;
    uint64_t next_lid = 0;
    // Manifest fields this engine does not use but carries from a loaded
    // snapshot to the next save (persistence.rs SnapshotManifest).
    uint32_t snap_m = 32, snap_ef_construction = 128, snap_max_elements = 0;
    // Parameters of the last fvh_build, reused by fvh_compact.
    uint32_t build_nlist = 0, build_coarse_iters = 25, build_pq_iters = 25;
    uint64_t build_seed = 0;
    // delta buffer
    uint32_t n_delta = 0, delta_cap = 0, delta_dead = 0;
    fvh::DevArray<float> delta_rows;
    fvh::DevArray<uint32_t> delta_ids, delta_alive;
    std::vector<uint32_t> h_delta_ids, h_delta_alive;
    std::unordered_map<uint32_t, uint32_t> lid2delta;
    // stats (metrics.rs)
    std::atomic<uint64_t> queries{0}, total_ns{0};

    ~FvEngine() {
        pool.clear();
        if (delta_stream) {
            (void)hipStreamSynchronize(delta_stream);
            (void)hipStreamDestroy(delta_stream);
        }
    }
};

namespace fvh {

static uint32_t words_for(uint32_t n) { return (n + 31) / 32; }

static hipError_t upload(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return hipSuccess;
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
}

// A caller mask is compacted with its count taken on the device (the scan
// then scores only the allowed rows); a mask of tombstones alone is tested
// per row.
static FvFlatSearchParams flat_params(const FvEngine* e, uint32_t k, const uint32_t* mask,
                                      bool user_mask) {
    FvFlatSearchParams p{};
    p.d_valid_mask = mask;
    p.valid_count = mask && user_mask ? FV_VALID_COUNT_DEVICE : 0u;
    p.metric = e->cfg.metric;
    p.top_k = k;
    p.shortlist_factor = e->cfg.shortlist_factor;
    p.flags = e->cfg.flags;
    return p;
}

// Sizes every stream's scratch for the loaded tables (largest query batch,
// largest k); called from the load / add paths, never from a search.
static hipError_t size_scratch(FvEngine* e) {
    const FvEngineConfig& c = e->cfg;
    const uint32_t B = c.max_batch, K = c.max_top_k;
    size_t ws = 0;
    if (e->mode & kFlat) {
        for (uint32_t nq = B; nq >= 1; nq /= 2) {
            for (int um = 0; um < 2; ++um) {
                FvFlatSearchParams p = flat_params(e, K, e->valid.p, um == 1);
                ws = std::max(ws, fv_flat_workspace_size(&e->flat, nq, &p));
            }
            if (nq == 1) break;
        }
        if (ws == 0) return hipErrorInvalidValue;
    }
    if (e->mode & kIvf) {
        size_t wi = 0;
        for (uint32_t nq = B; nq >= 1; nq /= 2) {
            wi = std::max(wi, fv_ivf_pq_workspace_size(&e->ivf, nq, c.ivf_nprobe, c.ivf_shortlist, K));
            if (nq == 1) break;
        }
        if (wi == 0) return hipErrorInvalidValue;
        ws = std::max(ws, wi);
    }
    size_t dws = 0;
    if (e->delta_cap)
        for (uint32_t nq = B; nq >= 1; nq /= 2) {
            dws = std::max(dws, fv_flat_exact_workspace_size(e->delta_cap, c.dim, nq, K,
                                                             (MetricType)c.metric));
            if (nq == 1) break;
        }
    const uint32_t umw = words_for(e->max_lid + 1);
    for (auto& sp : e->pool) {
        Scratch& s = *sp;
        hipError_t r;
        if (s.ws.n < ws && (r = s.ws.alloc(ws)) != hipSuccess) return r;
        if (s.dws.n < dws && (r = s.dws.alloc(dws)) != hipSuccess) return r;
        if (s.mask.n < words_for(e->n_main) && (r = s.mask.alloc(words_for(e->n_main))) != hipSuccess) return r;
        if (s.delta_mask.n < words_for(e->delta_cap) &&
            (r = s.delta_mask.alloc(words_for(e->delta_cap))) != hipSuccess)
            return r;
        if (s.user_mask.n < umw && (r = s.user_mask.alloc(umw)) != hipSuccess) return r;
        if (s.hmask.n < umw && (r = s.hmask.alloc(umw)) != hipSuccess) return r;
    }
    return hipSuccess;
}

static void track_lids(FvEngine* e, const uint32_t* lids, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t l = lids ? lids[i] : i;
        e->max_lid = std::max(e->max_lid, l);
        e->next_lid = std::max<uint64_t>(e->next_lid, (uint64_t)l + 1);
    }
}

static bool lids_unique(const uint32_t* lids, uint32_t n) {
    if (!lids) return true;
    std::unordered_set<uint32_t> seen;
    seen.reserve(n);
    for (uint32_t i = 0; i < n; ++i)
        if (!seen.insert(lids[i]).second) return false;
    return true;
}

// Slot of a logical id in the main index, or 0xFFFFFFFF.
static uint32_t main_slot(const FvEngine* e, uint32_t lid) {
    if (e->slot2lid.p) {
        auto it = e->lid2slot.find(lid);
        return it == e->lid2slot.end() ? 0xFFFFFFFFu : it->second;
    }
    return lid < e->n_main ? lid : 0xFFFFFFFFu;
}

// Delta alive bits from the tombstones (index.rs:696-714
// upload_delta_valid_mask): row i is alive unless its logical id is deleted.
static void rebuild_delta_alive(FvEngine* e) {
    e->h_delta_alive.assign(words_for(e->delta_cap), 0u);
    e->delta_dead = 0;
    for (uint32_t i = 0; i < e->n_delta; ++i) {
        if (e->tombstones.count(e->h_delta_ids[i])) ++e->delta_dead;
        else e->h_delta_alive[i >> 5] |= 1u << (i & 31u);
    }
}

// Resolved per-call options of one search.
struct CallPlan {
    int algo;           // kFlat or kIvf
    uint32_t nprobe;    // IVF-PQ
    uint32_t shortlist; // IVF-PQ
    int rerank;         // IVF-PQ
    uint32_t factor;    // flat shortlist factor
};

static int resolve(const FvEngine* e, const FvSearchOptions* o, uint32_t k, CallPlan& cp) {
    const FvEngineConfig& c = e->cfg;
    const uint32_t algo = o ? o->algo : FVH_ALGO_DEFAULT;
    if (algo == FVH_ALGO_DEFAULT) cp.algo = (e->mode & kFlat) ? kFlat : kIvf;
    else if (algo == FVH_ALGO_FLAT) cp.algo = kFlat;
    else if (algo == FVH_ALGO_IVFPQ) cp.algo = kIvf;
    else return (int)hipErrorInvalidValue;
    if (!(e->mode & cp.algo)) return -1;  // no such tables loaded
    cp.factor = o && o->shortlist_factor ? o->shortlist_factor : c.shortlist_factor;
    cp.nprobe = o && o->nprobe ? o->nprobe : c.ivf_nprobe;
    if (cp.algo == kIvf && cp.nprobe > e->ivf.nlist) cp.nprobe = e->ivf.nlist;
    if (o && o->rerank_k >= 0) {
        // ivf_pq_lookup.cu:351-362: S = (rerank_k == 0 ? k : rerank_k), at
        // least k, at most 10240; rerank unless rerank_k == 1.
        uint32_t S = o->rerank_k == 0 ? k : (uint32_t)o->rerank_k;
        if (S < k) S = k;
        if (S > 10240u) S = 10240u;
        cp.shortlist = S;
        cp.rerank = o->rerank_k != 1;
    } else {
        cp.shortlist = std::max(c.ivf_shortlist, k);
        cp.rerank = 1;
    }
    return 0;
}

// Grows the stream's workspace for a call shape beyond the engine's
// configuration (after the buffer's previous work has finished).
static hipError_t ensure_ws(Scratch& s, size_t need) {
    if (need <= s.ws.n) return hipSuccess;
    hipError_t he;
    if (s.used && (he = hipEventSynchronize(s.free_ev)) != hipSuccess) return he;
    return s.ws.alloc(need);
}

// One device pass of at most max_batch queries (index.rs:2180-2237 for the
// flat mode, :1979-2108 for IVF-PQ, :1353-1471 for the delta scan and merge).
static int search_pass(FvEngine* e, Scratch& s, hipStream_t st, const float* d_q, uint32_t nq,
                       uint32_t k, const uint32_t* d_user, uint32_t user_words,
                       GpuSearchResult* d_out, const CallPlan& cp) {
    hipError_t he;
    const FvEngineConfig& c = e->cfg;
    int rc;
    FvFlatSearchParams fp{};
    size_t need;
    if (cp.algo == kFlat) {
        fp = flat_params(e, k, nullptr, d_user != nullptr);
        fp.shortlist_factor = cp.factor;
        if (d_user || e->main_dead) fp.d_valid_mask = s.mask.p;
        need = fv_flat_workspace_size(&e->flat, nq, &fp);
    } else {
        need = fv_ivf_pq_workspace_size(&e->ivf, nq, cp.nprobe, cp.shortlist, k);
    }
    if (need == 0) return (int)hipErrorInvalidValue;
    if ((he = ensure_ws(s, need)) != hipSuccess) return (int)he;
    if (s.used && (he = hipStreamWaitEvent(st, s.free_ev, 0)) != hipSuccess) return (int)he;
    const bool has_delta = e->n_delta > e->delta_dead;
    GpuSearchResult* main_out = has_delta ? s.merge.p : d_out;
    GpuSearchResult* delta_out = s.merge.p + (size_t)nq * k;
    // Delta scan on the second stream, after the queries are in place.
    if (has_delta) {
        if ((he = hipEventRecord(s.queries_ready, st)) != hipSuccess) return (int)he;
        if ((he = hipStreamWaitEvent(e->delta_stream, s.queries_ready, 0)) != hipSuccess) return (int)he;
        const uint32_t* dmask = nullptr;
        if (d_user || e->delta_dead) {
            slot_mask_kernel<<<(e->n_delta + 255) / 256, 256, 0, e->delta_stream>>>(
                e->delta_alive.p, e->delta_ids.p, e->n_delta, d_user, user_words, s.delta_mask.p);
            dmask = s.delta_mask.p;
        }
        rc = fv_flat_exact_search(e->delta_rows.p, e->n_delta, c.dim, d_q, nq, k,
                                      (MetricType)c.metric, dmask, delta_out, s.dws.p, s.dws.n,
                                      e->delta_stream);
        if (rc) return rc;
        const size_t cnt = (size_t)nq * k;
        remap_ids_kernel<<<(unsigned)((cnt + 255) / 256), 256, 0, e->delta_stream>>>(
            delta_out, cnt, e->delta_ids.p, e->n_delta);
        if ((he = hipEventRecord(s.delta_done, e->delta_stream)) != hipSuccess) return (int)he;
    }
    // Main index.
    const uint32_t* mmask = nullptr;
    if (d_user || e->main_dead) {
        slot_mask_kernel<<<(e->n_main + 255) / 256, 256, 0, st>>>(
            e->valid.p, e->slot2lid.p, e->n_main, d_user, user_words, s.mask.p);
        mmask = s.mask.p;
    }
    if (cp.algo == kFlat) {
        rc = fv_flat_search(&e->flat, d_q, nq, &fp, main_out, s.ws.p, s.ws.n, st);
    } else {
        rc = fv_ivf_pq_search(&e->ivf, e->pq_codes_lo.p, d_q, nq, nullptr, k, nullptr,
                              cp.nprobe, cp.shortlist, cp.rerank, (MetricType)c.metric, mmask,
                              main_out, s.ws.p, s.ws.n, st);
    }
    if (rc) return rc;
    if (e->slot2lid.p) {
        const size_t cnt = (size_t)nq * k;
        remap_ids_kernel<<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(main_out, cnt,
                                                                      e->slot2lid.p, e->n_main);
    }
    if (has_delta) {
        if ((he = hipStreamWaitEvent(st, s.delta_done, 0)) != hipSuccess) return (int)he;
        if ((he = fv::launch_merge_topk(s.merge.p, 2, nq, k, k, d_out, st)) != hipSuccess) return (int)he;
    }
    if ((he = hipEventRecord(s.free_ev, st)) != hipSuccess) return (int)he;
    s.used = true;
    he = hipGetLastError();
    return he == hipSuccess ? 0 : (int)he;
}

static Scratch& pick(FvEngine* e) {
    const uint64_t i = e->rr.fetch_add(1, std::memory_order_relaxed);  // streams.rs:117-120
    return *e->pool[i % e->pool.size()];
}

static int check_search(const FvEngine* e, uint32_t top_k) {
    if (e->mode == 0) return -1;
    if (top_k == 0 || top_k > e->cfg.max_top_k) return (int)hipErrorInvalidValue;
    if (e->n_delta > e->delta_dead && 2ull * top_k > 4096ull) return (int)hipErrorInvalidValue;
    return 0;
}

}  // namespace fvh

extern "C" {

int fvh_create(const FvEngineConfig* cfg, FvEngine** out) {
    if (!cfg || !out) return -1;
    *out = nullptr;
    if (cfg->dim == 0 || cfg->metric > METRIC_INNER_PRODUCT || cfg->max_top_k > FV_MAX_TOPK)
        return (int)hipErrorInvalidValue;
    std::unique_ptr<FvEngine> e(new FvEngine());
    e->cfg = *cfg;
    FvEngineConfig& c = e->cfg;
    if (!c.num_streams) c.num_streams = 4;
    if (!c.max_batch) c.max_batch = 1024;
    if (!c.max_top_k) c.max_top_k = 256;
    if (!c.shortlist_factor) c.shortlist_factor = 8;
    if (!c.ivf_nprobe) c.ivf_nprobe = 32;
    if (!c.ivf_shortlist) c.ivf_shortlist = 1000;
    hipError_t he = hipStreamCreateWithFlags(&e->delta_stream, hipStreamNonBlocking);
    if (he != hipSuccess) return (int)he;
    for (uint32_t i = 0; i < c.num_streams; ++i) {
        std::unique_ptr<fvh::Scratch> s(new fvh::Scratch());
        if ((he = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess ||
            (he = hipEventCreateWithFlags(&s->queries_ready, hipEventDisableTiming)) != hipSuccess ||
            (he = hipEventCreateWithFlags(&s->delta_done, hipEventDisableTiming)) != hipSuccess ||
            (he = hipEventCreateWithFlags(&s->free_ev, hipEventDisableTiming)) != hipSuccess ||
            (he = s->q.alloc((size_t)c.max_batch * c.dim)) != hipSuccess ||
            (he = s->res.alloc((size_t)c.max_batch * c.max_top_k)) != hipSuccess ||
            (he = s->merge.alloc((size_t)2 * c.max_batch * c.max_top_k)) != hipSuccess ||
            (he = s->hq.alloc((size_t)c.max_batch * c.dim)) != hipSuccess ||
            (he = s->hres.alloc((size_t)c.max_batch * c.max_top_k)) != hipSuccess)
            return (int)he;
        e->pool.push_back(std::move(s));
    }
    *out = e.release();
    return 0;
}

void fvh_destroy(FvEngine* e) { delete e; }

static int set_ids(FvEngine* e, const uint32_t* lids, uint32_t n) {
    e->lid2slot.clear();
    e->h_slot2lid.clear();
    e->slot2lid.release();
    // A new main index starts with no tombstones and an empty delta buffer
    // (index.rs:519-524); the delta's device buffers keep their capacity.
    e->tombstones.clear();
    e->n_delta = e->delta_dead = 0;
    e->h_delta_ids.clear();
    e->h_delta_alive.clear();
    e->lid2delta.clear();
    e->max_lid = 0;
    e->next_lid = 0;
    if (lids) {
        e->h_slot2lid.assign(lids, lids + n);
        for (uint32_t s = 0; s < n; ++s) e->lid2slot[lids[s]] = s;
        hipError_t he = e->slot2lid.alloc(n);
        if (he == hipSuccess) he = fvh::upload(e->slot2lid.p, lids, (size_t)n * 4);
        if (he != hipSuccess) return (int)he;
    }
    fvh::track_lids(e, lids, n);
    e->h_valid.assign(fvh::words_for(n), 0xFFFFFFFFu);
    if (n % 32) e->h_valid.back() = (1u << (n % 32)) - 1u;
    e->main_dead = 0;
    hipError_t he = e->valid.alloc(e->h_valid.size());
    if (he == hipSuccess) he = fvh::upload(e->valid.p, e->h_valid.data(), e->h_valid.size() * 4);
    return (int)he;
}

}  // extern "C"

namespace fvh {

// Flat matrix-core tables encoded from the FP32 rows already on the device.
static int attach_flat(FvEngine* e, uint32_t dtype) {
    const uint32_t n = e->n_main, dim = e->cfg.dim;
    if (dtype != FV_DTYPE_F16 && dtype != FV_DTYPE_BF16 && dtype != FV_DTYPE_FP8_E4M3)
        return (int)hipErrorInvalidValue;
    const uint32_t step = dtype == FV_DTYPE_FP8_E4M3 ? 128u : 32u;
    const uint32_t pd = (dim + step - 1) / step * step;
    if (pd > (dtype == FV_DTYPE_FP8_E4M3 ? 3072u : 1536u)) return (int)hipErrorInvalidValue;
    const bool fp8 = dtype == FV_DTYPE_FP8_E4M3;
    const size_t esz = fp8 ? 1 : 2;
    hipError_t he;
    if ((he = e->codes.alloc((size_t)n * pd * esz)) != hipSuccess ||
        (he = e->norms.alloc(n)) != hipSuccess || (he = e->norms64.alloc(n)) != hipSuccess ||
        (he = e->scales.alloc(fp8 ? n : 0)) != hipSuccess ||
        (he = e->enc_err.alloc(fp8 ? 0 : n)) != hipSuccess)
        return (int)he;
    int rc = fv_encode_flat(e->rows.p, n, dim, pd, dtype, e->codes.p, e->norms.p, e->norms64.p,
                            e->scales.p, nullptr);
    if (rc == 0 && !fp8)
        rc = fv_encode_error(e->rows.p, e->codes.p, n, dim, pd, dtype, e->enc_err.p, nullptr);
    if (rc) return rc;
    if ((he = hipDeviceSynchronize()) != hipSuccess) return (int)he;
    e->flat = FvFlatTables{e->codes.p, e->scales.p, e->norms.p, e->norms64.p, e->rows.p, n, dim, pd,
                           dtype, fp8 ? nullptr : e->enc_err.p};
    e->mode |= kFlat;
    return 0;
}

static int alloc_ivf(FvEngine* e, uint32_t nlist, uint32_t m_pq) {
    const uint32_t n = e->n_main, dim = e->cfg.dim;
    hipError_t he;
    if ((he = e->centroids.alloc((size_t)nlist * dim)) != hipSuccess ||
        (he = e->codebooks.alloc((size_t)m_pq * 256 * (dim / m_pq))) != hipSuccess ||
        (he = e->pq_codes.alloc((size_t)n * m_pq)) != hipSuccess ||
        (he = e->pq_codes_lo.alloc((size_t)n * m_pq)) != hipSuccess ||
        (he = e->offsets.alloc(nlist + 1)) != hipSuccess || (he = e->vec_ids.alloc(n)) != hipSuccess)
        return (int)he;
    return 0;
}

// IVF-PQ tables in place (centroids ... vec_ids filled): list-ordered codes.
static int finish_ivf(FvEngine* e, uint32_t nlist, uint32_t m_pq) {
    const uint32_t dim = e->cfg.dim;
    e->ivf = IvfPqGpuTables{e->centroids.p, e->codebooks.p, e->pq_codes.p, e->offsets.p,
                            e->vec_ids.p, e->rows.p, e->n_main, dim, nlist, m_pq, dim / m_pq};
    int rc = fv_ivf_pq_list_order_codes(&e->ivf, e->pq_codes_lo.p, nullptr);
    if (rc) return rc;
    hipError_t he = hipDeviceSynchronize();
    if (he != hipSuccess) return (int)he;
    e->mode |= kIvf;
    return 0;
}

static int upload_rows(FvEngine* e, const float* h_vectors, uint32_t n) {
    hipError_t he;
    if ((he = e->rows.alloc((size_t)n * e->cfg.dim)) != hipSuccess ||
        (he = upload(e->rows.p, h_vectors, (size_t)n * e->cfg.dim * 4)) != hipSuccess)
        return (int)he;
    e->n_main = n;
    e->mode = 0;
    return 0;
}

}  // namespace fvh

extern "C" {

int fvh_load_flat(FvEngine* e, const float* h_vectors, uint32_t n, uint32_t dtype,
                  const uint32_t* h_logical_ids) {
    if (!e || !h_vectors) return -1;
    if (n == 0 || (dtype != FV_DTYPE_F16 && dtype != FV_DTYPE_BF16 && dtype != FV_DTYPE_FP8_E4M3))
        return (int)hipErrorInvalidValue;
    if (!fvh::lids_unique(h_logical_ids, n)) return (int)hipErrorInvalidValue;
    int rc;
    if ((rc = fvh::upload_rows(e, h_vectors, n)) || (rc = fvh::attach_flat(e, dtype)) ||
        (rc = set_ids(e, h_logical_ids, n)))
        return rc;
    return (int)fvh::size_scratch(e);
}

int fvh_load_ivfpq(FvEngine* e, const float* h_centroids, uint32_t nlist, const float* h_codebooks,
                   uint32_t m_pq, const uint8_t* h_codes, const uint32_t* h_offsets,
                   const uint32_t* h_vec_ids, const float* h_vectors, uint32_t n,
                   const uint32_t* h_logical_ids) {
    if (!e || !h_centroids || !h_codebooks || !h_codes || !h_offsets || !h_vec_ids || !h_vectors)
        return -1;
    const uint32_t dim = e->cfg.dim;
    if (n == 0 || nlist == 0 || m_pq == 0 || dim % m_pq) return (int)hipErrorInvalidValue;
    if (!fvh::lids_unique(h_logical_ids, n)) return (int)hipErrorInvalidValue;
    int rc;
    if ((rc = fvh::upload_rows(e, h_vectors, n)) || (rc = fvh::alloc_ivf(e, nlist, m_pq)))
        return rc;
    hipError_t he;
    if ((he = fvh::upload(e->centroids.p, h_centroids, (size_t)nlist * dim * 4)) != hipSuccess ||
        (he = fvh::upload(e->codebooks.p, h_codebooks, (size_t)m_pq * 256 * (dim / m_pq) * 4)) != hipSuccess ||
        (he = fvh::upload(e->pq_codes.p, h_codes, (size_t)n * m_pq)) != hipSuccess ||
        (he = fvh::upload(e->offsets.p, h_offsets, (size_t)(nlist + 1) * 4)) != hipSuccess ||
        (he = fvh::upload(e->vec_ids.p, h_vec_ids, (size_t)n * 4)) != hipSuccess)
        return (int)he;
    if ((rc = fvh::finish_ivf(e, nlist, m_pq)) || (rc = set_ids(e, h_logical_ids, n))) return rc;
    return (int)fvh::size_scratch(e);
}

}  // extern "C"

namespace fvh {

static int build_tables(FvEngine* e, const float* h_vectors, uint32_t n, const uint32_t* lids,
                        uint32_t flat_dtype, uint32_t nlist, uint32_t m_pq, uint32_t coarse_iters,
                        uint32_t pq_iters, uint64_t seed);

}  // namespace fvh

extern "C" {

int fvh_build(FvEngine* e, const float* h_vectors, uint32_t n, uint32_t flat_dtype,
              uint32_t nlist, uint32_t m_pq, uint32_t coarse_iters, uint32_t pq_iters,
              uint64_t seed) {
    if (!e || !h_vectors) return -1;
    return fvh::build_tables(e, h_vectors, n, nullptr, flat_dtype, nlist, m_pq, coarse_iters,
                             pq_iters, seed);
}

int fvh_compact(FvEngine* e) {
    using namespace fvh;
    if (!e) return -1;
    if (e->mode == 0) return -1;
    if (e->tombstones.empty() && e->n_delta == 0) return 0;  // index.rs:719-721
    const uint32_t dim = e->cfg.dim;
    hipError_t he = hipDeviceSynchronize();
    if (he != hipSuccess) return (int)he;
    // Live main rows in slot order, then live delta rows (index.rs:736-750).
    std::vector<float> main_rows((size_t)e->n_main * dim), delta_rows((size_t)e->n_delta * dim);
    if ((he = hipMemcpy(main_rows.data(), e->rows.p, main_rows.size() * 4,
                        hipMemcpyDeviceToHost)) != hipSuccess ||
        (e->n_delta && (he = hipMemcpy(delta_rows.data(), e->delta_rows.p,
                                       delta_rows.size() * 4, hipMemcpyDeviceToHost)) != hipSuccess))
        return (int)he;
    std::vector<float> rows;
    std::vector<uint32_t> lids;
    rows.reserve(main_rows.size() + delta_rows.size());
    for (uint32_t sl = 0; sl < e->n_main; ++sl) {
        const uint32_t lid = e->h_slot2lid.empty() ? sl : e->h_slot2lid[sl];
        if (e->tombstones.count(lid)) continue;
        lids.push_back(lid);
        rows.insert(rows.end(), main_rows.begin() + (size_t)sl * dim,
                    main_rows.begin() + (size_t)(sl + 1) * dim);
    }
    for (uint32_t j = 0; j < e->n_delta; ++j) {
        const uint32_t lid = e->h_delta_ids[j];
        if (e->tombstones.count(lid)) continue;
        lids.push_back(lid);
        rows.insert(rows.end(), delta_rows.begin() + (size_t)j * dim,
                    delta_rows.begin() + (size_t)(j + 1) * dim);
    }
    if (lids.empty()) return (int)hipErrorInvalidValue;  // nothing left to index
    const uint64_t watermark = e->next_lid;
    const uint32_t fd = (e->mode & kFlat) ? e->flat.dtype : 0u;
    const uint32_t nl = (e->mode & kIvf) ? std::max(e->build_nlist, e->ivf.nlist) : 0u;
    const uint32_t mp = (e->mode & kIvf) ? e->ivf.m_pq : 0u;
    const int rc = build_tables(e, rows.data(), (uint32_t)lids.size(), lids.data(), fd, nl, mp,
                                e->build_coarse_iters, e->build_pq_iters, e->build_seed);
    // index.rs:796-799: the id counter never goes back.
    e->next_lid = std::max(e->next_lid, watermark);
    return rc;
}

}  // extern "C"

namespace fvh {

static int build_tables(FvEngine* e, const float* h_vectors, uint32_t n, const uint32_t* lids,
                        uint32_t flat_dtype, uint32_t nlist, uint32_t m_pq, uint32_t coarse_iters,
                        uint32_t pq_iters, uint64_t seed) {
    const uint32_t dim = e->cfg.dim;
    if (n == 0 || (flat_dtype == 0 && (nlist == 0 || m_pq == 0))) return (int)hipErrorInvalidValue;
    if (m_pq && dim % m_pq) return (int)hipErrorInvalidValue;
    if (!lids_unique(lids, n)) return (int)hipErrorInvalidValue;
    e->build_nlist = nlist;
    e->build_coarse_iters = coarse_iters;
    e->build_pq_iters = pq_iters;
    e->build_seed = seed;
    int rc;
    if ((rc = fvh::upload_rows(e, h_vectors, n))) return rc;
    if (flat_dtype && (rc = fvh::attach_flat(e, flat_dtype))) return rc;
    if (nlist && m_pq) {
        const uint32_t nl = std::min(nlist, n);
        if ((rc = fvh::alloc_ivf(e, nl, m_pq))) return rc;
        const size_t wsb = fv_ivf_pq_build_workspace_size(n, dim, nl, m_pq);
        fvh::DevArray<uint8_t> ws;
        hipError_t he;
        if ((he = ws.alloc(std::max<size_t>(wsb, 1))) != hipSuccess) return (int)he;
        if ((rc = fv_ivf_pq_build(e->rows.p, n, dim, nl, m_pq, coarse_iters, pq_iters, seed,
                                  e->centroids.p, e->codebooks.p, e->pq_codes.p, e->offsets.p,
                                  e->vec_ids.p, nullptr, ws.p, ws.n, nullptr)))
            return rc;
        if ((rc = fvh::finish_ivf(e, nl, m_pq))) return rc;
    }
    if ((rc = set_ids(e, lids, n))) return rc;
    return (int)fvh::size_scratch(e);
}

}  // namespace fvh

extern "C" {

int fvh_info(const FvEngine* e, FvEngineInfo* out) {
    if (!e || !out) return -1;
    *out = FvEngineInfo{};
    out->tables = (uint32_t)e->mode;
    out->num_vectors = e->n_main;
    out->delta_len = e->n_delta;
    out->tombstones = (uint32_t)e->tombstones.size();
    out->next_logical_id = (uint32_t)std::min<uint64_t>(e->next_lid, UINT32_MAX);
    out->nlist = (e->mode & fvh::kIvf) ? e->ivf.nlist : 0u;
    out->m_pq = (e->mode & fvh::kIvf) ? e->ivf.m_pq : 0u;
    out->flat_dtype = (e->mode & fvh::kFlat) ? e->flat.dtype : 0u;
    return 0;
}

int fvh_add_delta(FvEngine* e, const float* h_vectors, const uint32_t* h_logical_ids, uint32_t n) {
    if (!e || !h_vectors || !h_logical_ids) return -1;
    if (e->mode == 0) return -1;
    if (n == 0) return 0;
    // Logical ids stay unique, as the reference's are (add_vectors hands out
    // next_logical_id, index.rs:667-669): an id already used by the main
    // index, the delta buffer or a deleted row is rejected, as is a repeat
    // inside the batch. Nothing is changed when the call fails.
    {
        std::unordered_set<uint32_t> batch;
        batch.reserve(n);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t lid = h_logical_ids[i];
            if (!batch.insert(lid).second || fvh::main_slot(e, lid) != 0xFFFFFFFFu ||
                e->lid2delta.count(lid) || e->tombstones.count(lid))
                return (int)hipErrorInvalidValue;
        }
    }
    const uint32_t dim = e->cfg.dim;
    if ((uint64_t)e->n_delta + n > UINT32_MAX / 2) return (int)hipErrorInvalidValue;
    const uint32_t need = e->n_delta + n;
    hipError_t he;
    if (need > e->delta_cap) {
        uint32_t cap = std::max(need, e->delta_cap * 2);
        fvh::DevArray<float> rows;
        fvh::DevArray<uint32_t> ids, alive;
        if ((he = rows.alloc((size_t)cap * dim)) != hipSuccess ||
            (he = ids.alloc(cap)) != hipSuccess ||
            (he = alive.alloc(fvh::words_for(cap))) != hipSuccess)
            return (int)he;
        if (e->n_delta &&
            (he = hipMemcpy(rows.p, e->delta_rows.p, (size_t)e->n_delta * dim * 4,
                            hipMemcpyDeviceToDevice)) != hipSuccess)
            return (int)he;
        std::swap(rows.p, e->delta_rows.p);
        std::swap(rows.n, e->delta_rows.n);
        std::swap(ids.p, e->delta_ids.p);
        std::swap(ids.n, e->delta_ids.n);
        std::swap(alive.p, e->delta_alive.p);
        std::swap(alive.n, e->delta_alive.n);
        e->delta_cap = cap;
    }
    if ((he = fvh::upload(e->delta_rows.p + (size_t)e->n_delta * dim, h_vectors,
                          (size_t)n * dim * 4)) != hipSuccess)
        return (int)he;
    for (uint32_t i = 0; i < n; ++i) {
        e->lid2delta[h_logical_ids[i]] = e->n_delta + i;
        e->h_delta_ids.push_back(h_logical_ids[i]);
    }
    e->n_delta = need;
    // Earlier rows keep their state: alive bits follow the tombstones.
    fvh::rebuild_delta_alive(e);
    fvh::track_lids(e, h_logical_ids, n);
    if ((he = fvh::upload(e->delta_ids.p, e->h_delta_ids.data(), (size_t)e->n_delta * 4)) != hipSuccess ||
        (he = fvh::upload(e->delta_alive.p, e->h_delta_alive.data(), e->h_delta_alive.size() * 4)) != hipSuccess)
        return (int)he;
    return (int)fvh::size_scratch(e);
}

int fvh_add_vectors(FvEngine* e, const float* h_vectors, uint32_t n, uint32_t* h_assigned_ids) {
    if (!e || !h_vectors || !h_assigned_ids) return -1;
    if (e->mode == 0) return -1;
    if (n == 0) return 0;
    if (e->next_lid + n > UINT32_MAX) return (int)hipErrorInvalidValue;
    for (uint32_t i = 0; i < n; ++i) h_assigned_ids[i] = (uint32_t)(e->next_lid + i);
    return fvh_add_delta(e, h_vectors, h_assigned_ids, n);
}

int fvh_delete(FvEngine* e, const uint32_t* h_logical_ids, uint32_t n) {
    if (!e || !h_logical_ids) return -1;
    // index.rs:583-640 delete_ids: an id already deleted is skipped; a delta
    // row is tombstoned; a main row is tombstoned and its validity bit cleared.
    bool delta_dirty = false, main_dirty = false;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t lid = h_logical_ids[i];
        if (e->tombstones.count(lid)) continue;
        if (e->lid2delta.count(lid)) {
            e->tombstones.insert(lid);
            delta_dirty = true;
            continue;
        }
        const uint32_t slot = fvh::main_slot(e, lid);
        if (slot != 0xFFFFFFFFu) {
            e->tombstones.insert(lid);
            if ((e->h_valid[slot >> 5] >> (slot & 31u)) & 1u) {
                e->h_valid[slot >> 5] &= ~(1u << (slot & 31u));
                ++e->main_dead;
            }
            main_dirty = true;
        }
    }
    hipError_t he = hipSuccess;
    if (main_dirty) he = fvh::upload(e->valid.p, e->h_valid.data(), e->h_valid.size() * 4);
    if (he == hipSuccess && delta_dirty) {
        fvh::rebuild_delta_alive(e);
        he = fvh::upload(e->delta_alive.p, e->h_delta_alive.data(), e->h_delta_alive.size() * 4);
    }
    return (int)he;
}

int fvh_next_logical_id(const FvEngine* e, uint32_t* out) {
    if (!e || !out) return -1;
    *out = (uint32_t)std::min<uint64_t>(e->next_lid, UINT32_MAX);
    return 0;
}

int fvh_search(FvEngine* e, const float* h_queries, uint32_t nq, uint32_t top_k,
               const uint32_t* h_mask_words, uint32_t mask_words, GpuSearchResult* h_out) {
    return fvh_search_ex(e, h_queries, nq, top_k, nullptr, h_mask_words, mask_words, h_out);
}

int fvh_search_ex(FvEngine* e, const float* h_queries, uint32_t nq, uint32_t top_k,
                  const FvSearchOptions* opts, const uint32_t* h_mask_words, uint32_t mask_words,
                  GpuSearchResult* h_out) {
    if (!e || !h_queries || !h_out) return -1;
    int rc = fvh::check_search(e, top_k);
    if (rc) return rc;
    fvh::CallPlan cp{};
    if ((rc = fvh::resolve(e, opts, top_k, cp))) return rc;
    if (nq == 0) return 0;
    const auto t0 = std::chrono::steady_clock::now();
    fvh::Scratch& s = fvh::pick(e);
    std::lock_guard<std::mutex> lock(s.mu);
    const uint32_t dim = e->cfg.dim, B = e->cfg.max_batch;
    hipError_t he;
    const uint32_t* d_user = nullptr;
    uint32_t uw = 0;
    if (h_mask_words) {
        uw = std::min<uint32_t>(mask_words, (uint32_t)s.user_mask.n);
        if (s.used && (he = hipEventSynchronize(s.free_ev)) != hipSuccess) return (int)he;
        std::memcpy(s.hmask.p, h_mask_words, (size_t)uw * 4);
        if (uw && (he = hipMemcpyAsync(s.user_mask.p, s.hmask.p, (size_t)uw * 4,
                                       hipMemcpyHostToDevice, s.stream)) != hipSuccess)
            return (int)he;
        d_user = s.user_mask.p;
    }
    for (uint32_t q0 = 0; q0 < nq; q0 += B) {
        const uint32_t nb = std::min(B, nq - q0);
        if (s.used && (he = hipEventSynchronize(s.free_ev)) != hipSuccess) return (int)he;
        std::memcpy(s.hq.p, h_queries + (size_t)q0 * dim, (size_t)nb * dim * 4);
        if ((he = hipMemcpyAsync(s.q.p, s.hq.p, (size_t)nb * dim * 4, hipMemcpyHostToDevice,
                                 s.stream)) != hipSuccess)
            return (int)he;
        if ((rc = fvh::search_pass(e, s, s.stream, s.q.p, nb, top_k, d_user, uw, s.res.p, cp)))
            return rc;
        if ((he = hipMemcpyAsync(s.hres.p, s.res.p, (size_t)nb * top_k * sizeof(GpuSearchResult),
                                 hipMemcpyDeviceToHost, s.stream)) != hipSuccess ||
            (he = hipStreamSynchronize(s.stream)) != hipSuccess)
            return (int)he;
        std::memcpy(h_out + (size_t)q0 * top_k, s.hres.p, (size_t)nb * top_k * sizeof(GpuSearchResult));
    }
    e->queries.fetch_add(nq, std::memory_order_relaxed);
    e->total_ns.fetch_add((uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                              std::chrono::steady_clock::now() - t0).count(),
                          std::memory_order_relaxed);
    return 0;
}

int fvh_search_device(FvEngine* e, const float* d_queries, uint32_t nq, uint32_t top_k,
                      const uint32_t* d_mask_words, uint32_t mask_words, GpuSearchResult* d_out,
                      void* stream) {
    return fvh_search_device_ex(e, d_queries, nq, top_k, nullptr, d_mask_words, mask_words, d_out,
                                stream);
}

int fvh_search_device_ex(FvEngine* e, const float* d_queries, uint32_t nq, uint32_t top_k,
                         const FvSearchOptions* opts, const uint32_t* d_mask_words,
                         uint32_t mask_words, GpuSearchResult* d_out, void* stream) {
    if (!e || !d_queries || !d_out) return -1;
    int rc = fvh::check_search(e, top_k);
    if (rc) return rc;
    fvh::CallPlan cp{};
    if ((rc = fvh::resolve(e, opts, top_k, cp))) return rc;
    if (nq == 0) return 0;
    const auto t0 = std::chrono::steady_clock::now();
    fvh::Scratch& s = fvh::pick(e);
    std::lock_guard<std::mutex> lock(s.mu);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t dim = e->cfg.dim, B = e->cfg.max_batch;
    for (uint32_t q0 = 0; q0 < nq; q0 += B) {
        const uint32_t nb = std::min(B, nq - q0);
        if ((rc = fvh::search_pass(e, s, st, d_queries + (size_t)q0 * dim, nb, top_k, d_mask_words,
                                   d_mask_words ? mask_words : 0u, d_out + (size_t)q0 * top_k, cp)))
            return rc;
    }
    e->queries.fetch_add(nq, std::memory_order_relaxed);
    e->total_ns.fetch_add((uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                              std::chrono::steady_clock::now() - t0).count(),
                          std::memory_order_relaxed);
    return 0;
}

int fvh_stats(const FvEngine* e, uint64_t* queries, double* total_ms) {
    if (!e) return -1;
    if (queries) *queries = e->queries.load();
    if (total_ms) *total_ms = (double)e->total_ns.load() / 1e6;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// FLSHVEC1 snapshots (crates/engine/src/persistence.rs:3-200,
// crates/engine/src/index.rs:845-1258).
// ---------------------------------------------------------------------------

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <map>
#include <string>

namespace fvh {

constexpr size_t kSnapAlign = 64;
constexpr char kSnapMagic[8] = {'F', 'L', 'S', 'H', 'V', 'E', 'C', '1'};

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

static size_t dtype_size(const std::string& t) {
    if (t == "f32" || t == "u32") return 4;
    if (t == "u16") return 2;
    if (t == "u8") return 1;
    return 0;
}

static void json_str(std::string& o, const char* s) {
    o += '"';
    for (; *s; ++s) {
        if (*s == '"' || *s == '\\') o += '\\';
        o += *s;
    }
    o += '"';
}

// Minimal JSON reader for the manifest: objects, arrays, strings, numbers
// (kept as text), true/false/null.
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    std::string text;  // number or string
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const char* k) const {
        for (const auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char* p;
    const char* end;
    int depth = 0;
    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool str(std::string& out) {
        if (p >= end || *p != '"') return false;
        ++p;
        while (p < end && *p != '"') {
            if (*p == '\\') {
                if (++p >= end) return false;
                const char c = *p;
                if (c == 'n') out += '\n';
                else if (c == 't') out += '\t';
                else if (c == 'r') out += '\r';
                else if (c == 'b') out += '\b';
                else if (c == 'f') out += '\f';
                else if (c == 'u') {
                    if (end - p < 5) return false;
                    out += '?';  // names here are ASCII; escapes are not interpreted
                    p += 4;
                } else out += c;
                ++p;
            } else {
                out += *p++;
            }
        }
        if (p >= end) return false;
        ++p;
        return true;
    }
    bool value(JVal& v) {
        if (++depth > 64) return false;
        ws();
        if (p >= end) return false;
        bool ok = true;
        if (*p == '{') {
            v.kind = JVal::Obj;
            ++p;
            ws();
            if (p < end && *p == '}') {
                ++p;
            } else {
                for (;;) {
                    ws();
                    std::string k;
                    JVal x;
                    if (!str(k)) { ok = false; break; }
                    ws();
                    if (p >= end || *p != ':') { ok = false; break; }
                    ++p;
                    if (!value(x)) { ok = false; break; }
                    v.obj.emplace_back(std::move(k), std::move(x));
                    ws();
                    if (p < end && *p == ',') { ++p; continue; }
                    if (p < end && *p == '}') { ++p; break; }
                    ok = false;
                    break;
                }
            }
        } else if (*p == '[') {
            v.kind = JVal::Arr;
            ++p;
            ws();
            if (p < end && *p == ']') {
                ++p;
            } else {
                for (;;) {
                    JVal x;
                    if (!value(x)) { ok = false; break; }
                    v.arr.push_back(std::move(x));
                    ws();
                    if (p < end && *p == ',') { ++p; continue; }
                    if (p < end && *p == ']') { ++p; break; }
                    ok = false;
                    break;
                }
            }
        } else if (*p == '"') {
            v.kind = JVal::Str;
            ok = str(v.text);
        } else if (*p == '-' || (*p >= '0' && *p <= '9')) {
            v.kind = JVal::Num;
            while (p < end && (*p == '-' || *p == '+' || *p == '.' || *p == 'e' || *p == 'E' ||
                               (*p >= '0' && *p <= '9')))
                v.text += *p++;
        } else if (end - p >= 4 && !strncmp(p, "true", 4)) {
            v.kind = JVal::Bool; v.text = "true"; p += 4;
        } else if (end - p >= 5 && !strncmp(p, "false", 5)) {
            v.kind = JVal::Bool; v.text = "false"; p += 5;
        } else if (end - p >= 4 && !strncmp(p, "null", 4)) {
            v.kind = JVal::Null; p += 4;
        } else {
            ok = false;
        }
        --depth;
        return ok;
    }
};

// Unsigned integer field; false when absent or not a plain unsigned integer.
static bool juint(const JVal* v, uint64_t& out) {
    if (!v || v->kind != JVal::Num || v->text.empty()) return false;
    uint64_t x = 0;
    for (char c : v->text) {
        if (c < '0' || c > '9') return false;
        const uint64_t d = (uint64_t)(c - '0');
        if (x > (UINT64_MAX - d) / 10) return false;
        x = x * 10 + d;
    }
    out = x;
    return true;
}

}  // namespace fvh

struct FvSnapshot {
    void* map = nullptr;
    size_t size = 0;
    size_t data_start = 0;
    struct Sec { size_t offset, byte_len, len; std::string dtype; };
    std::map<std::string, Sec> sections;
    ~FvSnapshot() {
        if (map) munmap(map, size);
    }
};

extern "C" {

int fvh_snapshot_write(const char* path, const FvSnapshotInfo* info,
                       const FvSnapshotSection* sections, uint32_t num_sections) {
    using namespace fvh;
    if (!path || !info || (num_sections && !sections)) return -1;
    struct Meta { size_t offset, byte_len; };
    std::map<std::string, std::pair<Meta, const FvSnapshotSection*>> sorted;
    std::vector<Meta> metas(num_sections);
    size_t cursor = 0;
    for (uint32_t i = 0; i < num_sections; ++i) {
        const FvSnapshotSection& s = sections[i];
        if (!s.name || !s.dtype) return -1;
        const size_t es = dtype_size(s.dtype);
        if (!es || (s.len && !s.data)) return FVH_ERR_FORMAT;
        cursor = align_up(cursor, kSnapAlign);
        metas[i] = Meta{cursor, (size_t)s.len * es};
        sorted[s.name] = {metas[i], &s};  // a repeated name keeps its last entry
        cursor += metas[i].byte_len;
    }
    // serde_json's compact form of SnapshotManifest, fields in declaration order.
    std::string j = "{\"version\":" + std::to_string(info->version) + ",\"config\":{\"dim\":" +
                    std::to_string(info->dim) + ",\"max_elements\":" +
                    std::to_string(info->max_elements) + ",\"m\":" + std::to_string(info->m) +
                    ",\"ef_construction\":" + std::to_string(info->ef_construction) +
                    ",\"nlist\":" + std::to_string(info->nlist) + ",\"m_pq\":" +
                    std::to_string(info->m_pq) + ",\"metric\":" + std::to_string(info->metric) +
                    "},\"num_vectors\":" + std::to_string(info->num_vectors) +
                    ",\"entry_point\":" + std::to_string(info->entry_point) +
                    ",\"cagra_degree\":" + std::to_string(info->cagra_degree) +
                    ",\"cagra_seeds\":[],\"rabitq_padded_dim\":" +
                    std::to_string(info->rabitq_padded_dim) + ",\"fp16_padded_dim\":" +
                    std::to_string(info->fp16_padded_dim) + ",\"sparse_num_docs\":" +
                    std::to_string(info->sparse_num_docs) + ",\"sparse_nnz\":" +
                    std::to_string(info->sparse_nnz) + ",\"next_logical_id\":" +
                    std::to_string(info->next_logical_id) + ",\"sections\":{";
    bool first = true;
    for (const auto& kv : sorted) {
        if (!first) j += ',';
        first = false;
        json_str(j, kv.first.c_str());
        j += ":{\"offset\":" + std::to_string(kv.second.first.offset) + ",\"byte_len\":" +
             std::to_string(kv.second.first.byte_len) + ",\"dtype\":";
        json_str(j, kv.second.second->dtype);
        j += ",\"len\":" + std::to_string(kv.second.second->len) + "}";
    }
    j += "}}";
    FILE* f = fopen(path, "wb");
    if (!f) return FVH_ERR_IO;
    bool ok = fwrite(kSnapMagic, 1, 8, f) == 8;
    uint8_t hl[8];
    for (int b = 0; b < 8; ++b) hl[b] = (uint8_t)((uint64_t)j.size() >> (8 * b));
    ok = ok && fwrite(hl, 1, 8, f) == 8 && fwrite(j.data(), 1, j.size(), f) == j.size();
    static const uint8_t zeros[kSnapAlign] = {0};
    const size_t header_end = 16 + j.size();
    const size_t pad = align_up(header_end, kSnapAlign) - header_end;
    ok = ok && (pad == 0 || fwrite(zeros, 1, pad, f) == pad);
    size_t written = 0;
    for (uint32_t i = 0; ok && i < num_sections; ++i) {
        const size_t gap = metas[i].offset - written;
        ok = gap == 0 || fwrite(zeros, 1, gap, f) == gap;
        if (ok && metas[i].byte_len)
            ok = fwrite(sections[i].data, 1, metas[i].byte_len, f) == metas[i].byte_len;
        written = metas[i].offset + metas[i].byte_len;
    }
    if (fclose(f) != 0) ok = false;
    return ok ? 0 : FVH_ERR_IO;
}

int fvh_snapshot_open(const char* path, FvSnapshot** out, FvSnapshotInfo* info) {
    using namespace fvh;
    if (!path || !out) return -1;
    *out = nullptr;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return FVH_ERR_IO;
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        return FVH_ERR_IO;
    }
    std::unique_ptr<FvSnapshot> s(new FvSnapshot());
    s->size = (size_t)st.st_size;
    if (s->size < 16) {
        close(fd);
        return FVH_ERR_FORMAT;
    }
    s->map = mmap(nullptr, s->size, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (s->map == MAP_FAILED) {
        s->map = nullptr;
        return FVH_ERR_IO;
    }
    const uint8_t* b = static_cast<const uint8_t*>(s->map);
    if (memcmp(b, kSnapMagic, 8) != 0) return FVH_ERR_FORMAT;
    uint64_t hlen = 0;
    for (int i = 0; i < 8; ++i) hlen |= (uint64_t)b[8 + i] << (8 * i);
    if (hlen > s->size - 16) return FVH_ERR_FORMAT;   // truncated header
    const size_t header_end = 16 + (size_t)hlen;
    JParser jp{reinterpret_cast<const char*>(b + 16), reinterpret_cast<const char*>(b + header_end)};
    JVal root;
    if (!jp.value(root) || root.kind != JVal::Obj) return FVH_ERR_FORMAT;
    jp.ws();
    if (jp.p != jp.end) return FVH_ERR_FORMAT;
    s->data_start = align_up(header_end, kSnapAlign);
    if (s->data_start > s->size) return FVH_ERR_FORMAT;  // truncated payload
    FvSnapshotInfo I{};
    const JVal* cfg = root.get("config");
    const JVal* secs = root.get("sections");
    uint64_t v = 0;
    auto u32f = [&](const JVal* o, const char* k, uint32_t& dst) {
        if (!juint(o ? o->get(k) : nullptr, v) || v > UINT32_MAX) return false;
        dst = (uint32_t)v;
        return true;
    };
    auto u64f = [&](const char* k, uint64_t& dst) { return juint(root.get(k), dst); };
    const JVal* seeds = root.get("cagra_seeds");
    if (!cfg || cfg->kind != JVal::Obj || !secs || secs->kind != JVal::Obj || !seeds ||
        seeds->kind != JVal::Arr || !u32f(&root, "version", I.version) ||
        !u32f(cfg, "dim", I.dim) || !u32f(cfg, "max_elements", I.max_elements) ||
        !u32f(cfg, "m", I.m) || !u32f(cfg, "ef_construction", I.ef_construction) ||
        !u32f(cfg, "nlist", I.nlist) || !u32f(cfg, "m_pq", I.m_pq) ||
        !u32f(cfg, "metric", I.metric) || !u64f("num_vectors", I.num_vectors) ||
        !u32f(&root, "entry_point", I.entry_point) || !u64f("cagra_degree", I.cagra_degree) ||
        !u64f("rabitq_padded_dim", I.rabitq_padded_dim) ||
        !u64f("fp16_padded_dim", I.fp16_padded_dim) ||
        !u64f("sparse_num_docs", I.sparse_num_docs) || !u64f("sparse_nnz", I.sparse_nnz) ||
        !u32f(&root, "next_logical_id", I.next_logical_id))
        return FVH_ERR_FORMAT;
    I.num_cagra_seeds = (uint32_t)seeds->arr.size();
    for (const auto& kv : secs->obj) {
        if (kv.second.kind != JVal::Obj) return FVH_ERR_FORMAT;
        FvSnapshot::Sec sec;
        uint64_t off, bl, len;
        const JVal* dt = kv.second.get("dtype");
        if (!juint(kv.second.get("offset"), off) || !juint(kv.second.get("byte_len"), bl) ||
            !juint(kv.second.get("len"), len) || !dt || dt->kind != JVal::Str)
            return FVH_ERR_FORMAT;
        // Every section must lie inside the file (persistence.rs:150-158).
        if (off > s->size || bl > s->size || s->data_start + off > s->size - bl)
            return FVH_ERR_FORMAT;
        sec.offset = (size_t)off;
        sec.byte_len = (size_t)bl;
        sec.len = (size_t)len;
        sec.dtype = dt->text;
        s->sections[kv.first] = sec;
    }
    if (info) *info = I;
    *out = s.release();
    return 0;
}

int fvh_snapshot_section(const FvSnapshot* s, const char* name, const void** data,
                         uint64_t* byte_len, uint64_t* len, char* dtype8) {
    if (!s || !name) return -1;
    auto it = s->sections.find(name);
    if (it == s->sections.end()) return 1;
    const auto& sec = it->second;
    if (data) *data = static_cast<const uint8_t*>(s->map) + s->data_start + sec.offset;
    if (byte_len) *byte_len = sec.byte_len;
    if (len) *len = sec.len;
    if (dtype8) {
        const size_t n = std::min<size_t>(sec.dtype.size(), 7);
        memcpy(dtype8, sec.dtype.data(), n);
        dtype8[n] = 0;
    }
    return 0;
}

void fvh_snapshot_close(FvSnapshot* s) { delete s; }

}  // extern "C"

namespace fvh {

template <typename T>
static hipError_t download(std::vector<T>& dst, const T* src, size_t count) {
    dst.resize(count);
    if (count == 0) return hipSuccess;
    return hipMemcpy(dst.data(), src, count * sizeof(T), hipMemcpyDeviceToHost);
}

// Typed view of a section: its bytes must be a whole number of T and the
// pointer aligned (persistence.rs:173-200); absent sections give nullptr.
template <typename T>
static int section_view(const FvSnapshot* s, const char* name, const T** p, size_t* count) {
    const void* d = nullptr;
    uint64_t bl = 0;
    *p = nullptr;
    *count = 0;
    const int rc = fvh_snapshot_section(s, name, &d, &bl, nullptr, nullptr);
    if (rc == 1) return 1;
    if (rc) return rc;
    if (bl % sizeof(T) || (reinterpret_cast<uintptr_t>(d) % alignof(T))) return FVH_ERR_FORMAT;
    *p = static_cast<const T*>(d);
    *count = (size_t)(bl / sizeof(T));
    return 0;
}

}  // namespace fvh

extern "C" {

int fvh_save_snapshot(FvEngine* e, const char* path) {
    using namespace fvh;
    if (!e || !path) return -1;
    if (e->mode == 0) return -1;
    const uint32_t dim = e->cfg.dim, n = e->n_main;
    hipError_t he = hipDeviceSynchronize();
    if (he != hipSuccess) return (int)he;
    std::vector<float> rows, centroids, codebooks, norms, scales, delta_rows;
    std::vector<uint8_t> codes, pq_codes;
    std::vector<uint32_t> offsets, vec_ids, lids(n), tomb;
    std::vector<uint16_t> f16;
    if ((he = download(rows, e->rows.p, (size_t)n * dim)) != hipSuccess) return (int)he;
    for (uint32_t s = 0; s < n; ++s) lids[s] = e->h_slot2lid.empty() ? s : e->h_slot2lid[s];
    FvSnapshotInfo I{};
    I.version = 1;
    I.dim = dim;
    I.max_elements = std::max(n, e->snap_max_elements);
    I.m = e->snap_m;
    I.ef_construction = e->snap_ef_construction;
    I.metric = e->cfg.metric;
    I.num_vectors = n;
    uint32_t nlist = 0, m_pq = 0;
    const char* extra_name = nullptr;
    if (e->mode & kIvf) {
        nlist = e->ivf.nlist;
        m_pq = e->ivf.m_pq;
        if ((he = download(centroids, e->centroids.p, (size_t)nlist * dim)) != hipSuccess ||
            (he = download(codebooks, e->codebooks.p, (size_t)m_pq * 256 * (dim / m_pq))) != hipSuccess ||
            (he = download(pq_codes, e->pq_codes.p, (size_t)n * m_pq)) != hipSuccess ||
            (he = download(offsets, e->offsets.p, (size_t)nlist + 1)) != hipSuccess ||
            (he = download(vec_ids, e->vec_ids.p, (size_t)n)) != hipSuccess)
            return (int)he;
    }
    if (e->mode & kFlat) {
        const FvFlatTables& t = e->flat;
        const size_t esz = t.dtype == FV_DTYPE_FP8_E4M3 ? 1 : 2;
        if ((he = download(codes, static_cast<const uint8_t*>(t.d_codes), (size_t)n * t.padded_dim * esz)) != hipSuccess ||
            (he = download(norms, e->norms.p, (size_t)n)) != hipSuccess)
            return (int)he;
        if (t.dtype == FV_DTYPE_F16) {
            // The reference pads FP16 rows to a multiple of 16 (quant.rs:306-327);
            // the padding is zero either way.
            const uint32_t pd16 = (dim + 15) / 16 * 16;
            f16.resize((size_t)n * pd16);
            const uint16_t* src = reinterpret_cast<const uint16_t*>(codes.data());
            for (uint32_t r = 0; r < n; ++r)
                memcpy(&f16[(size_t)r * pd16], src + (size_t)r * t.padded_dim, pd16 * 2);
            I.fp16_padded_dim = pd16;
        } else {
            extra_name = t.dtype == FV_DTYPE_BF16 ? "vectors_bf16" : "vectors_fp8";
            if (t.dtype == FV_DTYPE_FP8_E4M3 &&
                (he = download(scales, e->scales.p, (size_t)n)) != hipSuccess)
                return (int)he;
        }
    }
    I.nlist = nlist;
    I.m_pq = m_pq;
    // Tombstones ascending (RoaringBitset::to_vec, index.rs:984-991).
    tomb.assign(e->tombstones.begin(), e->tombstones.end());
    std::sort(tomb.begin(), tomb.end());
    // next_logical_id: the engine's counter, which already covers every id
    // it holds and the value of a loaded snapshot (ids the reference handed
    // out and compacted away are never reused).
    uint64_t next = e->next_lid;
    for (uint32_t l : lids) next = std::max<uint64_t>(next, (uint64_t)l + 1);
    for (uint32_t l : e->h_delta_ids) next = std::max<uint64_t>(next, (uint64_t)l + 1);
    I.next_logical_id = (uint32_t)std::min<uint64_t>(next, UINT32_MAX);
    if (e->n_delta &&
        (he = download(delta_rows, e->delta_rows.p, (size_t)e->n_delta * dim)) != hipSuccess)
        return (int)he;
    // index.rs:866-1028 section order.
    std::vector<FvSnapshotSection> S = {
        {"vectors", "f32", rows.size(), rows.data()},
        {"logical_ids", "u32", lids.size(), lids.data()},
        {"main_valid_words", "u32", e->h_valid.size(), e->h_valid.data()},
        {"hnsw_adjacency", "u32", 0, nullptr},
        {"hnsw_degrees", "u32", 0, nullptr},
        {"cagra_adjacency", "u32", 0, nullptr},
        {"ivf_centroids", "f32", centroids.size(), centroids.data()},
        {"pq_codebooks", "f32", codebooks.size(), codebooks.data()},
        {"pq_codes", "u8", pq_codes.size(), pq_codes.data()},
        {"ivf_offsets", "u32", offsets.size(), offsets.data()},
        {"ivf_vec_ids", "u32", vec_ids.size(), vec_ids.data()},
        {"sq_mins", "f32", 0, nullptr},
        {"sq_scales", "f32", 0, nullptr},
        {"sq_codes", "u8", 0, nullptr},
        {"rabitq_signs", "f32", 0, nullptr},
        {"rabitq_centroids", "f32", 0, nullptr},
        {"rabitq_bits", "u32", 0, nullptr},
        {"rabitq_meta", "f32", 0, nullptr},
        {"vectors_fp16", "u16", f16.size(), f16.data()},
        {"vec_norms_sq", "f32", norms.size(), norms.data()},
    };
    if (extra_name) {
        const bool is8 = e->flat.dtype == FV_DTYPE_FP8_E4M3;
        S.push_back({extra_name, is8 ? "u8" : "u16", is8 ? codes.size() : codes.size() / 2,
                     codes.data()});
        if (is8) S.push_back({"row_scales", "f32", scales.size(), scales.data()});
    }
    if (!tomb.empty()) S.push_back({"tombstones", "u32", tomb.size(), tomb.data()});
    if (e->n_delta) {
        S.push_back({"delta_vectors", "f32", delta_rows.size(), delta_rows.data()});
        S.push_back({"delta_ids", "u32", e->h_delta_ids.size(), e->h_delta_ids.data()});
    }
    return fvh_snapshot_write(path, &I, S.data(), (uint32_t)S.size());
}

int fvh_load_snapshot(FvEngine* e, const char* path, uint32_t flat_dtype) {
    using namespace fvh;
    if (!e || !path) return -1;
    FvSnapshot* raw = nullptr;
    FvSnapshotInfo I{};
    int rc = fvh_snapshot_open(path, &raw, &I);
    if (rc) return rc;
    std::unique_ptr<FvSnapshot> s(raw);
    const uint32_t dim = e->cfg.dim;
    if (I.dim != dim || I.metric != e->cfg.metric) return (int)hipErrorInvalidValue;
    const uint64_t n64 = I.num_vectors;
    if (n64 == 0 || n64 > UINT32_MAX) return FVH_ERR_FORMAT;
    const uint32_t n = (uint32_t)n64;
    const float* rows; const uint32_t* lids; const uint32_t* valid;
    size_t c_rows, c_lids, c_valid;
    if (section_view(s.get(), "vectors", &rows, &c_rows) ||
        section_view(s.get(), "logical_ids", &lids, &c_lids) ||
        section_view(s.get(), "main_valid_words", &valid, &c_valid))
        return FVH_ERR_FORMAT;
    if (c_rows != (size_t)n * dim || c_lids != n || c_valid < words_for(n)) return FVH_ERR_FORMAT;
    // Identity ids load as "no remap" (index.rs:1187-1190 id_remapped).
    bool remapped = false;
    for (uint32_t sl = 0; sl < n && !remapped; ++sl) remapped = lids[sl] != sl;
    const uint32_t* use_lids = remapped ? lids : nullptr;
    // Low byte: FvDtype of flat tables encoded from the rows (0 = none);
    // FVH_LOAD_IVFPQ: the file's IVF-PQ tables (also implied by 0).
    const uint32_t dtype = flat_dtype & 0xFFu;
    const bool want_ivf = dtype == 0 || (flat_dtype & FVH_LOAD_IVFPQ);
    if ((flat_dtype & ~(0xFFu | FVH_LOAD_IVFPQ)) != 0) return (int)hipErrorInvalidValue;
    if (want_ivf) {
        const float* cent; const float* books; const uint8_t* codes; const uint32_t* off;
        const uint32_t* vids;
        size_t c_cent, c_books, c_codes, c_off, c_vids;
        if (section_view(s.get(), "ivf_centroids", &cent, &c_cent) ||
            section_view(s.get(), "pq_codebooks", &books, &c_books) ||
            section_view(s.get(), "pq_codes", &codes, &c_codes) ||
            section_view(s.get(), "ivf_offsets", &off, &c_off) ||
            section_view(s.get(), "ivf_vec_ids", &vids, &c_vids))
            return FVH_ERR_FORMAT;
        if (c_off < 2 || c_cent != (c_off - 1) * dim || c_vids != n || c_codes % n ||
            c_codes == 0)
            return FVH_ERR_FORMAT;
        const uint32_t nlist = (uint32_t)(c_off - 1), m_pq = (uint32_t)(c_codes / n);
        if (m_pq == 0 || dim % m_pq || c_books != (size_t)m_pq * 256 * (dim / m_pq) ||
            off[nlist] != n)
            return FVH_ERR_FORMAT;
        for (size_t i = 0; i < c_vids; ++i)
            if (vids[i] >= n) return FVH_ERR_FORMAT;
        for (uint32_t c = 0; c < nlist; ++c)
            if (off[c] > off[c + 1]) return FVH_ERR_FORMAT;
        rc = fvh_load_ivfpq(e, cent, nlist, books, m_pq, codes, off, vids, rows, n, use_lids);
        if (rc == 0 && dtype) {
            rc = attach_flat(e, dtype);
            if (rc == 0) rc = (int)size_scratch(e);
        }
    } else {
        rc = fvh_load_flat(e, rows, n, dtype, use_lids);
    }
    if (rc) return rc;
    const float* drows; const uint32_t* dids;
    size_t c_drows, c_dids;
    const int a = section_view(s.get(), "delta_vectors", &drows, &c_drows);
    const int b = section_view(s.get(), "delta_ids", &dids, &c_dids);
    if ((a && a != 1) || (b && b != 1)) return FVH_ERR_FORMAT;
    if (a == 0 && b == 0 && c_dids) {
        if (c_drows != c_dids * dim) return FVH_ERR_FORMAT;
        if ((rc = fvh_add_delta(e, drows, dids, (uint32_t)c_dids))) return rc;
    }
    const uint32_t* tomb;
    size_t c_tomb;
    const int t = section_view(s.get(), "tombstones", &tomb, &c_tomb);
    if (t == 0) {
        if (c_tomb) rc = fvh_delete(e, tomb, (uint32_t)c_tomb);
        // index.rs:1195-1198 keeps every listed id, also ids no row holds.
        for (size_t i = 0; i < c_tomb; ++i) e->tombstones.insert(tomb[i]);
    } else if (t == 1) {
        // index.rs:1199-1206: without tombstones, cleared validity bits.
        std::vector<uint32_t> dead;
        for (uint32_t sl = 0; sl < n; ++sl)
            if (!((valid[sl >> 5] >> (sl & 31u)) & 1u)) dead.push_back(lids[sl]);
        if (!dead.empty()) rc = fvh_delete(e, dead.data(), (uint32_t)dead.size());
    } else {
        return FVH_ERR_FORMAT;
    }
    if (rc) return rc;
    // The manifest's id counter and the fields this engine only carries.
    e->next_lid = std::max<uint64_t>(e->next_lid, I.next_logical_id);
    e->snap_m = I.m;
    e->snap_ef_construction = I.ef_construction;
    e->snap_max_elements = (uint32_t)std::min<uint64_t>(I.max_elements, UINT32_MAX);
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Micro-batcher (crates/server/src/batcher.rs:80-262): one worker thread
// coalesces single-query submissions into batched searches.
// ---------------------------------------------------------------------------

#include <condition_variable>
#include <deque>
#include <system_error>
#include <thread>

namespace fvh {

struct BatchRequest {
    const float* query;  // caller's buffer; valid until its submit returns
    uint32_t top_k;
    GpuSearchResult* out;
    double latency_us = 0.0;
    int rc = 0;
    bool done = false;
};

}  // namespace fvh

struct FvBatcher {
    FvBatchExecutor exec = nullptr;
    void* user = nullptr;
    uint32_t dim = 0;
    uint32_t max_batch = 1;
    std::chrono::microseconds max_wait{0};
    std::mutex mu;
    std::condition_variable work_cv;  // queue not empty, or stopping
    std::condition_variable done_cv;  // some request finished
    std::deque<fvh::BatchRequest*> queue;
    bool stop = false;
    std::atomic<uint64_t> total_requests{0};
    std::atomic<uint64_t> total_batches{0};
    std::thread worker;
};

namespace fvh {

// Runs the queries of one launch key as one batch and fills their results.
static void run_group(FvBatcher* b, const std::vector<BatchRequest*>& group,
                      std::vector<float>& qbuf, std::vector<GpuSearchResult>& rbuf) {
    const uint32_t nq = (uint32_t)group.size(), k = group[0]->top_k, dim = b->dim;
    qbuf.resize((size_t)nq * dim);
    rbuf.resize((size_t)nq * k);
    for (uint32_t i = 0; i < nq; ++i)
        std::memcpy(qbuf.data() + (size_t)i * dim, group[i]->query, (size_t)dim * sizeof(float));
    b->total_requests.fetch_add(nq, std::memory_order_relaxed);
    b->total_batches.fetch_add(1, std::memory_order_relaxed);
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = b->exec(b->user, k, qbuf.data(), nq, rbuf.data());
    const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    for (uint32_t i = 0; i < nq; ++i) {
        if (rc == 0)
            std::memcpy(group[i]->out, rbuf.data() + (size_t)i * k, (size_t)k * sizeof(GpuSearchResult));
        group[i]->rc = rc;
        group[i]->latency_us = us;
    }
}

static void batcher_loop(FvBatcher* b) {
    std::vector<BatchRequest*> items;
    std::vector<std::pair<uint32_t, std::vector<BatchRequest*>>> groups;
    std::vector<float> qbuf;
    std::vector<GpuSearchResult> rbuf;
    std::unique_lock<std::mutex> lk(b->mu);
    for (;;) {
        b->work_cv.wait(lk, [b] { return b->stop || !b->queue.empty(); });
        if (b->queue.empty()) return;  // stopping and drained
        items.clear();
        items.push_back(b->queue.front());
        b->queue.pop_front();
        const auto deadline = std::chrono::steady_clock::now() + b->max_wait;
        while (items.size() < b->max_batch) {
            if (!b->queue.empty()) {
                items.push_back(b->queue.front());
                b->queue.pop_front();
                continue;
            }
            if (b->stop || std::chrono::steady_clock::now() >= deadline) break;
            b->work_cv.wait_until(lk, deadline);
        }
        lk.unlock();
        // Partition by launch key, groups in the order of their first query.
        groups.clear();
        for (BatchRequest* r : items) {
            auto it = std::find_if(groups.begin(), groups.end(),
                                   [r](const auto& g) { return g.first == r->top_k; });
            if (it == groups.end()) groups.push_back({r->top_k, {r}});
            else it->second.push_back(r);
        }
        for (const auto& g : groups) {
            run_group(b, g.second, qbuf, rbuf);
            lk.lock();
            for (BatchRequest* r : g.second) r->done = true;
            lk.unlock();
            b->done_cv.notify_all();
        }
        lk.lock();
    }
}

static int engine_executor(void* user, uint32_t top_k, const float* queries, uint32_t nq,
                           GpuSearchResult* out) {
    return fvh_search(static_cast<FvEngine*>(user), queries, nq, top_k, nullptr, 0, out);
}

}  // namespace fvh

extern "C" {

int fvh_batcher_create_with_executor(FvBatchExecutor exec, void* user, uint32_t dim,
                                     uint32_t max_wait_us, uint32_t max_batch, FvBatcher** out) {
    if (!exec || !out) return -1;
    *out = nullptr;
    if (dim == 0) return (int)hipErrorInvalidValue;
    auto b = std::make_unique<FvBatcher>();
    b->exec = exec;
    b->user = user;
    b->dim = dim;
    b->max_batch = max_batch ? max_batch : 1u;  // batcher.rs:136
    b->max_wait = std::chrono::microseconds(max_wait_us);
    try {
        b->worker = std::thread(fvh::batcher_loop, b.get());
    } catch (const std::system_error&) {
        return (int)hipErrorOutOfMemory;  // no thread could be started
    }
    *out = b.release();
    return 0;
}

int fvh_batcher_create(FvEngine* e, uint32_t max_wait_us, uint32_t max_batch, FvBatcher** out) {
    if (!e || !out) return -1;
    return fvh_batcher_create_with_executor(fvh::engine_executor, e, e->cfg.dim, max_wait_us,
                                            max_batch, out);
}

int fvh_batcher_submit(FvBatcher* b, const float* h_query, uint32_t top_k, GpuSearchResult* h_out,
                       double* latency_us) {
    if (!b || !h_query || !h_out) return -1;
    if (top_k == 0) return (int)hipErrorInvalidValue;
    fvh::BatchRequest r;
    r.query = h_query;
    r.top_k = top_k;
    r.out = h_out;
    std::unique_lock<std::mutex> lk(b->mu);
    if (b->stop) return FVH_ERR_SHUTDOWN;
    b->queue.push_back(&r);
    b->work_cv.notify_one();
    b->done_cv.wait(lk, [&r] { return r.done; });
    if (latency_us) *latency_us = r.latency_us;
    return r.rc;
}

int fvh_batcher_stats(const FvBatcher* b, uint64_t* total_requests, uint64_t* total_batches,
                      double* avg_batch_size) {
    if (!b) return -1;
    const uint64_t nr = b->total_requests.load(std::memory_order_relaxed);
    const uint64_t nb = b->total_batches.load(std::memory_order_relaxed);
    if (total_requests) *total_requests = nr;
    if (total_batches) *total_batches = nb;
    if (avg_batch_size) *avg_batch_size = nb ? (double)nr / (double)nb : 0.0;
    return 0;
}

void fvh_batcher_destroy(FvBatcher* b) {
    if (!b) return;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        b->stop = true;
    }
    b->work_cv.notify_all();
    if (b->worker.joinable()) b->worker.join();
    delete b;
}

}  // extern "C"
