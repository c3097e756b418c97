// Library-owned device scratch for the reference's entry points, which take no
// workspace argument (cuda_flat_scan_batch, cuda_fp16_tc_search_batch,
// cuda_ivf_pq_search_batch). Instead of allocating and freeing their scratch
// on every call (the reference does cudaMallocAsync / cudaFreeAsync per call
// in its IVF stage 2, kernels/src/ivf_pq_lookup.cu:397-443), each (device,
// stream) keeps one buffer that grows to the largest call and is reused, as
// the reference host's per-stream scratchpads are (crates/engine/src/
// streams.rs:43-69).
//
// Safety: a lease holds its entry's mutex from acquisition until the caller's
// kernels are enqueued, so two host threads sharing a stream cannot resize
// the buffer under each other; work on one stream runs in order, so the next
// call's kernels never overlap the previous call's. An event recorded after
// each use lets an entry be freed (on growth or eviction) only once its last
// work has finished, also when its stream has been destroyed since.
#include <atomic>
#include <memory>
#include <mutex>
#include <vector>

#include "fv_workspace.hpp"

namespace fv {

namespace {

struct Entry {
    int device = -1;
    hipStream_t stream = nullptr;
    void* p = nullptr;
    size_t bytes = 0;
    hipEvent_t last = nullptr;  // recorded after the last use
    bool pending = false;       // `last` has been recorded
    uint64_t used = 0;
    std::atomic<int> users{0};  // leases held or pending (taken under g_mu)
    std::mutex mu;
};

constexpr size_t kMaxEntries = 16;
std::mutex g_mu;
std::vector<std::unique_ptr<Entry>> g_entries;
uint64_t g_tick = 0;

// Frees an entry's buffer once its last work has finished (entry lock held).
hipError_t release_buffer(Entry& e) {
    hipError_t r = hipSuccess;
    if (e.pending && e.last) r = hipEventSynchronize(e.last);
    e.pending = false;
    if (e.p) {
        const hipError_t f = hipFree(e.p);
        if (r == hipSuccess) r = f;
    }
    e.p = nullptr;
    e.bytes = 0;
    return r;
}

}  // namespace

WorkspaceLease::~WorkspaceLease() { release(); }

void WorkspaceLease::release() {
    if (!entry_) return;
    Entry* e = static_cast<Entry*>(entry_);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(e->stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone &&
        hipEventRecord(e->last, e->stream) == hipSuccess)
        e->pending = true;
    entry_ = nullptr;
    ptr = nullptr;
    e->mu.unlock();
    e->users.fetch_sub(1);  // last access: the entry may be evicted from here on
}

hipError_t stream_workspace(hipStream_t stream, size_t bytes, WorkspaceLease& lease) {
    lease.release();
    int dev = 0;
    hipError_t r = hipGetDevice(&dev);
    if (r != hipSuccess) return r;
    Entry* e = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        for (auto& x : g_entries)
            if (x->device == dev && x->stream == stream) e = x.get();
        if (!e) {
            if (g_entries.size() >= kMaxEntries) {
                // Evict the least recently used entry that no call holds or
                // waits for (users are counted under g_mu, so none can appear).
                size_t victim = g_entries.size();
                for (size_t i = 0; i < g_entries.size(); ++i) {
                    if (g_entries[i]->users.load() != 0) continue;
                    if (victim == g_entries.size() || g_entries[i]->used < g_entries[victim]->used)
                        victim = i;
                }
                if (victim < g_entries.size()) {
                    Entry& v = *g_entries[victim];
                    (void)release_buffer(v);
                    if (v.last) (void)hipEventDestroy(v.last);
                    g_entries.erase(g_entries.begin() + victim);
                }
            }
            auto ne = std::make_unique<Entry>();
            ne->device = dev;
            ne->stream = stream;
            if ((r = hipEventCreateWithFlags(&ne->last, hipEventDisableTiming)) != hipSuccess)
                return r;
            e = ne.get();
            g_entries.push_back(std::move(ne));
        }
        e->used = ++g_tick;
        e->users.fetch_add(1);
    }
    e->mu.lock();
    if (e->bytes < bytes) {
        // Grow with 25% headroom so that slowly growing calls do not
        // reallocate every time.
        if ((r = release_buffer(*e)) != hipSuccess) {
            e->mu.unlock();
            e->users.fetch_sub(1);
            return r;
        }
        const size_t want = bytes + bytes / 4;
        if ((r = hipMalloc(&e->p, want)) != hipSuccess) {
            e->p = nullptr;
            e->mu.unlock();
            e->users.fetch_sub(1);
            return r;
        }
        e->bytes = want;
    } else if (e->pending) {
        // The same handle may name a new stream (the old one destroyed with
        // work in flight): order this call after the buffer's last use.
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone &&
            (r = hipStreamWaitEvent(stream, e->last, 0)) != hipSuccess) {
            e->mu.unlock();
            e->users.fetch_sub(1);
            return r;
        }
    }
    lease.entry_ = e;
    lease.ptr = e->p;
    return hipSuccess;
}

}  // namespace fv
