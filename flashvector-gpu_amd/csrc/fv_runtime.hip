// Device, memory and stream entry points of the C ABI on HIP, plus the two
// known-answer-test kernels.
//
// Reference: kernels/src/hnsw_traverse.cu:525-637 (runtime wrappers on the
// CUDA runtime) and :387-437,694-733 (KAT kernels). Return codes keep the
// reference's convention: 0 success, -1 for a NULL required pointer, otherwise
// the runtime's error code.
#include "fv_flat_exact.hpp"

namespace fv {

__global__ void fill_results_kernel(GpuSearchResult* out, size_t count, uint32_t id_bits,
                                    uint32_t dist_bits) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        GpuSearchResult r;
        r.id = id_bits;
        r.distance = __uint_as_float(dist_bits);
        out[i] = r;
    }
}

__global__ void fill_words_kernel(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t stream) {
    if (n) fill_words_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(p, n, v);
}

hipError_t fill_results(GpuSearchResult* out, size_t count, uint32_t id_bits, uint32_t dist_bits,
                        hipStream_t stream) {
    if (count == 0) return hipSuccess;
    size_t blocks = (count + 255) / 256;
    fill_results_kernel<<<(unsigned)blocks, 256, 0, stream>>>(out, count, id_bits, dist_bits);
    return hipGetLastError();
}

// Dense distance matrix, one thread per (query, vector) pair, exact FP32 in the
// oracle's order (the reference used one warp per pair; the values it was
// checked against are the same sequential sums, cuda_sanity_test.rs:88-99).
__global__ void distance_matrix_kernel(const float* __restrict__ q, const float* __restrict__ x,
                                       uint32_t nq, uint32_t n, uint32_t dim, int metric,
                                       float* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nq * n) return;
    uint32_t qi = (uint32_t)(i / n), vi = (uint32_t)(i % n);
    out[i] = exact_distance_rt(q + (size_t)qi * dim, x + (size_t)vi * dim, dim, metric);
}

// Register bitonic sort of up to 64 (key, value) pairs across one wavefront:
// each compare-exchange partner is fetched with a cross-lane xor shuffle.
__global__ void wave_bitonic_test_kernel(const float* __restrict__ kin, const uint32_t* __restrict__ vin,
                                         float* __restrict__ kout, uint32_t* __restrict__ vout,
                                         uint32_t n) {
    const int lane = threadIdx.x;
    float key = (uint32_t)lane < n ? kin[lane] : 1e30f;
    uint32_t val = (uint32_t)lane < n ? vin[lane] : 0xFFFFFFFFu;
    for (int k = 2; k <= kWave; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            float ok = __shfl_xor(key, j, kWave);
            uint32_t ov = __shfl_xor(val, j, kWave);
            bool lower = (lane & j) == 0;
            bool up = (lane & k) == 0;
            // (key, value) lexicographic so that equal keys order deterministically.
            bool other_less = ok < key || (ok == key && ov < val);
            bool take = (lower == up) ? other_less : !other_less && !(ok == key && ov == val);
            if (take) {
                key = ok;
                val = ov;
            }
        }
    }
    if ((uint32_t)lane < n) {
        kout[lane] = key;
        vout[lane] = val;
    }
}

}  // namespace fv

extern "C" {

int cuda_init_device(int device_id) {
    hipError_t e = hipSetDevice(device_id);
    if (e != hipSuccess) return (int)e;
    return (int)hipFree(nullptr);
}

int cuda_set_device(int device_id) { return (int)hipSetDevice(device_id); }

int cuda_get_device_count(int* count) {
    if (!count) return -1;
    return (int)hipGetDeviceCount(count);
}

int cuda_get_device_memory(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return -1;
    return (int)hipMemGetInfo(free_bytes, total_bytes);
}

int cuda_device_synchronize(void) { return (int)hipDeviceSynchronize(); }

int cuda_enable_peer_access(int device, int peer) {
    if (device == peer) return 0;
    int can = 0;
    hipError_t e = hipDeviceCanAccessPeer(&can, device, peer);
    if (e != hipSuccess) return (int)e;
    if (!can) return (int)hipErrorPeerAccessUnsupported;
    int prev = 0;
    (void)hipGetDevice(&prev);
    e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    e = hipDeviceEnablePeerAccess(peer, 0);
    if (e == hipErrorPeerAccessAlreadyEnabled) {
        (void)hipGetLastError();
        e = hipSuccess;
    }
    (void)hipSetDevice(prev);
    return (int)e;
}

int cuda_memcpy_peer_async(void* dst, int dst_device, const void* src, int src_device,
                           size_t bytes, void* stream) {
    if (!dst || !src) return -1;
    if (bytes == 0) return 0;
    return (int)hipMemcpyPeerAsync(dst, dst_device, src, src_device, bytes,
                                   static_cast<hipStream_t>(stream));
}

int cuda_malloc_device(void** ptr, size_t bytes) {
    if (!ptr || bytes == 0) return -1;
    return (int)hipMalloc(ptr, bytes);
}

int cuda_free_device(void* ptr) {
    if (!ptr) return 0;
    return (int)hipFree(ptr);
}

// Mapped pinned host memory (cudaHostAllocMapped in the reference).
int cuda_malloc_host(void** ptr, size_t bytes) {
    if (!ptr || bytes == 0) return -1;
    return (int)hipHostMalloc(ptr, bytes, hipHostMallocMapped);
}

int cuda_free_host(void* ptr) {
    if (!ptr) return 0;
    return (int)hipHostFree(ptr);
}

int cuda_host_get_device_pointer(void** dev_ptr, void* host_ptr) {
    if (!dev_ptr || !host_ptr) return -1;
    return (int)hipHostGetDevicePointer(dev_ptr, host_ptr, 0);
}

int cuda_memcpy_h2d_async(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src) return -1;
    if (bytes == 0) return 0;
    return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                               static_cast<hipStream_t>(stream));
}

int cuda_memcpy_d2h_async(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src) return -1;
    if (bytes == 0) return 0;
    return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                               static_cast<hipStream_t>(stream));
}

int cuda_memset_device_async(void* dst, int value, size_t bytes, void* stream) {
    if (!dst) return -1;
    if (bytes == 0) return 0;
    return (int)hipMemsetAsync(dst, value, bytes, static_cast<hipStream_t>(stream));
}

// Non-blocking stream at the default priority.
int cuda_create_stream(void** stream) {
    if (!stream) return -1;
    return (int)hipStreamCreateWithFlags(reinterpret_cast<hipStream_t*>(stream),
                                         hipStreamNonBlocking);
}

int cuda_destroy_stream(void* stream) {
    if (!stream) return 0;
    return (int)hipStreamDestroy(static_cast<hipStream_t>(stream));
}

int cuda_sync_stream(void* stream) {
    if (!stream) return (int)hipDeviceSynchronize();
    return (int)hipStreamSynchronize(static_cast<hipStream_t>(stream));
}

int cuda_compute_distances_warp(const float* d_queries, const float* d_dataset,
                                uint32_t num_queries, uint32_t num_vectors, uint32_t dim,
                                MetricType metric, float* d_out_distances, void* stream) {
    if (!d_queries || !d_dataset || !d_out_distances) return -1;
    size_t total = (size_t)num_queries * num_vectors;
    if (total == 0) return 0;
    fv::distance_matrix_kernel<<<(unsigned)((total + 255) / 256), 256, 0,
                                 static_cast<hipStream_t>(stream)>>>(
        d_queries, d_dataset, num_queries, num_vectors, dim, (int)metric, d_out_distances);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

// Sorts up to one wavefront (64) of pairs ascending; the reference sorted 32.
int cuda_bitonic_sort_test(const float* d_keys_in, const uint32_t* d_vals_in, float* d_keys_out,
                           uint32_t* d_vals_out, uint32_t n, void* stream) {
    if (!d_keys_in || !d_vals_in || !d_keys_out || !d_vals_out) return -1;
    if (n > (uint32_t)fv::kWave) return (int)hipErrorInvalidValue;
    fv::wave_bitonic_test_kernel<<<1, fv::kWave, 0, static_cast<hipStream_t>(stream)>>>(
        d_keys_in, d_vals_in, d_keys_out, d_vals_out, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Kernel timer: brackets the dominant kernel of each search call with two HIP
// events on the call's stream, so that a caller can read that kernel's device
// time without a profiler.
// ---------------------------------------------------------------------------
#include <mutex>
#include <vector>

namespace fv {
namespace {
std::mutex g_timer_mu;
bool g_timer_on = false;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_timer_pairs;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_timer_free;
}  // namespace

void* timer_begin(hipStream_t stream) {
    std::lock_guard<std::mutex> g(g_timer_mu);
    if (!g_timer_on) return nullptr;
    std::pair<hipEvent_t, hipEvent_t> p;
    if (!g_timer_free.empty()) {
        p = g_timer_free.back();
        g_timer_free.pop_back();
    } else {
        if (hipEventCreate(&p.first) != hipSuccess) return nullptr;
        if (hipEventCreate(&p.second) != hipSuccess) return nullptr;
    }
    if (hipEventRecord(p.first, stream) != hipSuccess) return nullptr;
    g_timer_pairs.push_back(p);
    return reinterpret_cast<void*>(p.second);
}

void timer_end(void* token, hipStream_t stream) {
    if (token) (void)hipEventRecord(reinterpret_cast<hipEvent_t>(token), stream);
}

}  // namespace fv

extern "C" int fv_kernel_timer_enable(int enable) {
    std::lock_guard<std::mutex> g(fv::g_timer_mu);
    fv::g_timer_on = enable != 0;
    return 0;
}

extern "C" int fv_kernel_timer_read(double* total_ms, uint32_t* launches, int reset) {
    if (!total_ms || !launches) return -1;
    std::lock_guard<std::mutex> g(fv::g_timer_mu);
    double sum = 0.0;
    for (auto& p : fv::g_timer_pairs) {
        hipError_t e = hipEventSynchronize(p.second);
        if (e != hipSuccess) return (int)e;
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, p.first, p.second);
        if (e != hipSuccess) return (int)e;
        sum += ms;
    }
    *total_ms = sum;
    *launches = (uint32_t)fv::g_timer_pairs.size();
    if (reset) {
        for (auto& p : fv::g_timer_pairs) fv::g_timer_free.push_back(p);
        fv::g_timer_pairs.clear();
    }
    return 0;
}
