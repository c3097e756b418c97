// IVF-PQ index construction on the GPU (fv_build.h): k-means for the coarse
// quantizer and for every PQ sub-space, residual encoding and posting lists.
//
// Reference: crates/engine/src/cluster.rs:28-287 (KMeans::fit,
// ProductQuantizer::train / encode) and crates/engine/src/index.rs:405-458
// (their use in GpuVectorIndex::build), all on the host with rayon there.
//
// Here:
//  * km_assign_kernel: 64 rows x 64 centroids per tile, 4x4 register tiles of
//    exact distances in the reference's order (sequential over d, non-fused),
//    the running best (distance, centroid) key per row = first strict minimum.
//  * The update needs every cluster's rows in row order (the reference adds
//    rows into their centroid in index order): a stable radix sort of
//    (cluster, row) pairs (hipCUB) gives members and offsets, which are also
//    the posting lists; km_update_kernel sums each (cluster, dim) sequentially
//    over its members, km_finalize_kernel scales by 1/count and measures moves.
//  * k-means++: the pool rows' D^2 to the nearest chosen centroid are updated
//    per step in parallel; one workgroup sums them in f64 over fixed chunks and
//    locates the target. Draws (pool shuffle, targets, empty-cluster recovery)
//    come from a host splitmix64 stream in a fixed order.
#include <hipcub/hipcub.hpp>

#include <cfloat>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/flashvector/fv_build.h"
#include "fv_flat_exact.hpp"

namespace fv {
namespace {

constexpr int kAR = 64;         // rows per assign workgroup
constexpr int kAC = 64;         // centroids per tile
constexpr int kAStride = 68;    // padded LDS row (floats)
constexpr uint32_t kSeedChunk = 256;
constexpr uint32_t kEmptyDraws = 500;

__device__ __forceinline__ float km_l2(const float* __restrict__ a, const float* __restrict__ b,
                                       uint32_t dim) {
    float acc = 0.0f;
    for (uint32_t d = 0; d < dim; ++d) {
        float t = a[d] - b[d];
        float sq = t * t;
        acc = acc + sq;
    }
    return acc;
}

// DK dims per LDS stage (8, 16 or 32): sub-space rows of 8 dims stage no
// padding.
template <int DK>
__global__ __launch_bounds__(256) void km_assign_kernel(
    const float* __restrict__ X, uint32_t n, uint32_t dim, uint32_t ld,
    const float* __restrict__ C, uint32_t k, uint32_t* __restrict__ assign,
    uint8_t* __restrict__ codes, uint32_t code_stride) {
    constexpr int kADK = DK;
    constexpr int kRowsPerPass = 256 / (DK / 4);  // rows staged per pass of the loop below
    __shared__ __attribute__((aligned(16))) float sx[kADK * kAStride];
    __shared__ __attribute__((aligned(16))) float sc[kADK * kAStride];
    __shared__ uint64_t s_best[kAR][17];
    const int tid = threadIdx.x;
    const uint32_t r0 = blockIdx.x * kAR;
    const int tr = (tid & 15) * 4;   // this thread's 4 rows
    const int tc = (tid >> 4) * 4;   // and 4 centroid columns
    uint64_t best[4] = {kEmptyKey, kEmptyKey, kEmptyKey, kEmptyKey};

    for (uint32_t c0 = 0; c0 < k; c0 += kAC) {
        float acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
        for (uint32_t d0 = 0; d0 < dim; d0 += kADK) {
            __syncthreads();
            for (int it = 0; it * kRowsPerPass < kAR; ++it) {
                const int qq = tid / (DK / 4) + it * kRowsPerPass;
                const int dd = (tid % (DK / 4)) * 4;
                if (qq >= kAR) continue;
                const uint32_t rg = r0 + qq, cg = c0 + qq;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const uint32_t dg = d0 + dd + c;
                    sx[(dd + c) * kAStride + qq] =
                        (rg < n && dg < dim) ? X[(size_t)rg * ld + dg] : 0.0f;
                    sc[(dd + c) * kAStride + qq] =
                        (cg < k && dg < dim) ? C[(size_t)cg * dim + dg] : 0.0f;
                }
            }
            __syncthreads();
            const int dlim = min((int)kADK, (int)(dim - d0));
            for (int dd = 0; dd < dlim; ++dd) {
                const float4 xv = *reinterpret_cast<const float4*>(&sx[dd * kAStride + tr]);
                const float4 cv = *reinterpret_cast<const float4*>(&sc[dd * kAStride + tc]);
                const float xa[4] = {xv.x, xv.y, xv.z, xv.w};
                const float ca[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        float t = xa[i] - ca[j];
                        float sq = t * t;
                        acc[i][j] = acc[i][j] + sq;
                    }
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t c = c0 + tc + j;
            if (c >= k) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // dist < +inf: an infinite or NaN distance never becomes best.
                if (acc[i][j] < __builtin_inff()) {
                    const uint64_t key = make_key(acc[i][j], c);
                    best[i] = key < best[i] ? key : best[i];
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) s_best[tr + i][tid >> 4] = best[i];
    __syncthreads();
    if (tid < kAR) {
        uint64_t b = kEmptyKey;
        for (int j = 0; j < 16; ++j) b = s_best[tid][j] < b ? s_best[tid][j] : b;
        const uint32_t r = r0 + tid;
        if (r < n) {
            const uint32_t a = b == kEmptyKey ? 0u : key_slot(b);
            if (assign) assign[r] = a;
            if (codes) codes[(size_t)r * code_stride] = (uint8_t)a;
        }
    }
}

__global__ void km_iota_kernel(uint32_t* __restrict__ v, uint32_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

// offsets[c] = first position of cluster c in the sorted keys (n past the
// end); every c in [0, k] is written by exactly one thread.
__global__ void km_offsets_kernel(const uint32_t* __restrict__ skeys, uint32_t n, uint32_t k,
                                  uint32_t* __restrict__ offsets) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const uint32_t cur = i < n ? skeys[i] : k;
    if (i == 0) {
        for (uint32_t c = 0; c <= cur; ++c) offsets[c] = 0;
    } else {
        const uint32_t prev = skeys[i - 1];
        for (uint32_t c = prev + 1; c <= cur; ++c) offsets[c] = (uint32_t)i;
    }
}

// Sum of each (cluster, dim) over the cluster's rows in row order, from 0.
// The adds form one dependent chain per (cluster, dim); kChainLoads row
// loads are issued ahead of their adds so the chain waits on memory once per
// group rather than once per row.
constexpr int kChainLoads = 32;
__global__ __launch_bounds__(64) void km_update_kernel(
    const float* __restrict__ X, uint32_t dim, uint32_t ld, const uint32_t* __restrict__ members,
    const uint32_t* __restrict__ offsets, float* __restrict__ sums) {
    const uint32_t c = blockIdx.x;
    const uint32_t d = blockIdx.y * 64u + threadIdx.x;
    if (d >= dim) return;
    const uint32_t beg = offsets[c], end = offsets[c + 1];
    float acc = 0.0f;
    uint32_t j = beg;
    for (; j + kChainLoads <= end; j += kChainLoads) {
        float v[kChainLoads];
#pragma unroll
        for (int u = 0; u < kChainLoads; ++u) v[u] = X[(size_t)members[j + u] * ld + d];
#pragma unroll
        for (int u = 0; u < kChainLoads; ++u) acc = acc + v[u];
    }
    for (; j < end; ++j) acc = acc + X[(size_t)members[j] * ld + d];
    sums[(size_t)c * dim + d] = acc;
}

// New centroids of the non-empty clusters (sum * (1.0f / count)), their
// squared moves (sequential over d) and the largest move (non-negative float
// bits compare as unsigned).
__global__ void km_finalize_kernel(const float* __restrict__ sums, const float* __restrict__ cent,
                                   const uint32_t* __restrict__ offsets, uint32_t k, uint32_t dim,
                                   float* __restrict__ newc, uint32_t* __restrict__ sizes,
                                   uint32_t* __restrict__ max_bits) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k) return;
    const uint32_t cnt = offsets[c + 1] - offsets[c];
    sizes[c] = cnt;
    if (cnt == 0) return;
    const float inv = 1.0f / (float)cnt;
    float delta = 0.0f;
    for (uint32_t d = 0; d < dim; ++d) {
        const float v = sums[(size_t)c * dim + d] * inv;
        const float diff = v - cent[(size_t)c * dim + d];
        const float sq = diff * diff;
        delta = delta + sq;
        newc[(size_t)c * dim + d] = v;
    }
    atomicMax(max_bits, __float_as_uint(delta));
}

__global__ void km_copy_row_kernel(const float* __restrict__ X, uint32_t ld, const uint32_t* __restrict__ row,
                                   uint32_t dim, float* __restrict__ dst) {
    const uint32_t r = *row;
    for (uint32_t d = threadIdx.x; d < dim; d += blockDim.x) dst[d] = X[(size_t)r * ld + d];
}

// k-means++ step, part 1: D^2 of every pool row to its nearest chosen
// centroid (strict improvement only) and the f64 sum of each 256-row chunk in
// row order.
__global__ __launch_bounds__(256) void km_seed_update_kernel(
    const float* __restrict__ X, uint32_t dim, uint32_t ld, const uint32_t* __restrict__ pool_ix,
    uint32_t pool, const float* __restrict__ last, float* __restrict__ closest,
    double* __restrict__ chunk_sum) {
    __shared__ float s_v[kSeedChunk];
    const uint32_t i = blockIdx.x * kSeedChunk + threadIdx.x;
    float v = 0.0f;
    if (i < pool) {
        const float cand = km_l2(X + (size_t)pool_ix[i] * ld, last, dim);
        float cur = closest[i];
        if (cand < cur) {
            cur = cand;
            closest[i] = cand;
        }
        v = cur;
    }
    s_v[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t m = min(kSeedChunk, pool - blockIdx.x * kSeedChunk);
        double s = 0.0;
        for (uint32_t t = 0; t < m; ++t) s = s + (double)s_v[t];
        chunk_sum[blockIdx.x] = s;
    }
}

// k-means++ step, part 2 (one workgroup): total = chunk sums in order; the
// chosen row is the first whose running sum exceeds u * total (the pool's last
// row when total is 0 or rounding leaves none); its row becomes centroid c.
// The sums run in one thread over values the workgroup staged in LDS.
constexpr uint32_t kPickTile = 1024;
__global__ __launch_bounds__(256) void km_seed_pick_kernel(
    const double* __restrict__ chunk_sum, uint32_t nch, const float* __restrict__ closest,
    uint32_t pool, const uint32_t* __restrict__ pool_ix, const double* __restrict__ u,
    const float* __restrict__ X, uint32_t dim, uint32_t ld, float* __restrict__ dst) {
    __shared__ double s_cs[kPickTile];
    __shared__ float s_cl[kSeedChunk];
    __shared__ double s_run;
    __shared__ int s_hit;  // chunk holding the target, -1 while none
    __shared__ uint32_t s_row;
    const uint32_t tid = threadIdx.x;
    // Pass 1: total over all chunk sums, tile by tile in order.
    if (tid == 0) s_run = 0.0;
    for (uint32_t t0 = 0; t0 < nch; t0 += kPickTile) {
        const uint32_t m = min(kPickTile, nch - t0);
        __syncthreads();
        for (uint32_t j = tid; j < m; j += blockDim.x) s_cs[j] = chunk_sum[t0 + j];
        __syncthreads();
        if (tid == 0) {
            double run = s_run;
            for (uint32_t j = 0; j < m; ++j) run = run + s_cs[j];
            s_run = run;
        }
    }
    __syncthreads();
    const double total = s_run;
    uint32_t chosen = pool_ix[pool - 1];
    __syncthreads();  // every thread has read s_run before it is reset
    if (total > 0.0) {
        const double target = (*u) * total;
        // Pass 2: the first chunk whose running sum exceeds the target.
        if (tid == 0) {
            s_run = 0.0;
            s_hit = -1;
        }
        double before = 0.0;
        for (uint32_t t0 = 0; t0 < nch; t0 += kPickTile) {
            const uint32_t m = min(kPickTile, nch - t0);
            __syncthreads();
            if (s_hit >= 0) break;
            for (uint32_t j = tid; j < m; j += blockDim.x) s_cs[j] = chunk_sum[t0 + j];
            __syncthreads();
            if (tid == 0) {
                double run = s_run;
                for (uint32_t j = 0; j < m; ++j) {
                    if (run + s_cs[j] > target) {
                        s_hit = (int)(t0 + j);
                        break;
                    }
                    run = run + s_cs[j];
                }
                s_run = run;
            }
        }
        __syncthreads();
        const int hit = s_hit;
        before = s_run;
        if (hit >= 0) {
            // Pass 3: the first row of that chunk whose running sum exceeds it.
            const uint32_t b = (uint32_t)hit * kSeedChunk, e = min(b + kSeedChunk, pool);
            for (uint32_t i = b + tid; i < e; i += blockDim.x) s_cl[i - b] = closest[i];
            __syncthreads();
            if (tid == 0) {
                uint32_t pick = pool_ix[e - 1];
                double r = before;
                for (uint32_t i = b; i < e; ++i) {
                    r = r + (double)s_cl[i - b];
                    if (r > target) {
                        pick = pool_ix[i];
                        break;
                    }
                }
                s_row = pick;
            }
            __syncthreads();
            chosen = s_row;
        }
    }
    for (uint32_t d = tid; d < dim; d += blockDim.x) dst[d] = X[(size_t)chosen * ld + d];
}

// Mean of all rows per dim: sequential over rows, then times 1.0f / n.
__global__ void km_mean_kernel(const float* __restrict__ X, uint32_t n, uint32_t dim, uint32_t ld,
                               float* __restrict__ mean) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= dim) return;
    float acc = 0.0f;
    uint32_t i = 0;
    for (; i + kChainLoads <= n; i += kChainLoads) {
        float v[kChainLoads];
#pragma unroll
        for (int u = 0; u < kChainLoads; ++u) v[u] = X[(size_t)(i + u) * ld + d];
#pragma unroll
        for (int u = 0; u < kChainLoads; ++u) acc = acc + v[u];
    }
    for (; i < n; ++i) acc = acc + X[(size_t)i * ld + d];
    mean[d] = acc * (1.0f / (float)n);
}

// Empty-cluster recovery, one workgroup per empty cluster: of `ndraw` drawn
// rows the first strictly furthest from the mean (start: the extra first
// draw at distance 0) becomes the centroid.
__global__ __launch_bounds__(256) void km_empty_kernel(
    const float* __restrict__ X, uint32_t dim, uint32_t ld, const float* __restrict__ mean,
    const uint32_t* __restrict__ empty_ids, const uint32_t* __restrict__ draws, uint32_t ndraw,
    float* __restrict__ newc) {
    __shared__ float s_d[kEmptyDraws];
    __shared__ uint32_t s_far;
    const uint32_t e = blockIdx.x;
    const uint32_t* dr = draws + (size_t)e * (ndraw + 1);
    for (uint32_t t = threadIdx.x; t < ndraw; t += blockDim.x)
        s_d[t] = km_l2(X + (size_t)dr[1 + t] * ld, mean, dim);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t far = dr[0];
        float max_d = 0.0f;
        for (uint32_t t = 0; t < ndraw; ++t)
            if (s_d[t] > max_d) {
                max_d = s_d[t];
                far = dr[1 + t];
            }
        s_far = far;
    }
    __syncthreads();
    float* dst = newc + (size_t)empty_ids[e] * dim;
    for (uint32_t d = threadIdx.x; d < dim; d += blockDim.x) dst[d] = X[(size_t)s_far * ld + d];
}

__global__ void km_residual_kernel(const float* __restrict__ X, const float* __restrict__ C,
                                   const uint32_t* __restrict__ assign, uint32_t n, uint32_t dim,
                                   float* __restrict__ res) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * dim) return;
    const size_t r = i / dim, d = i % dim;
    res[i] = X[i] - C[(size_t)assign[r] * dim + d];
}

__global__ void km_fill_f32_kernel(float* __restrict__ p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- host side -----------------------------------------------------------

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    uint32_t below(uint32_t m) { return (uint32_t)(next() % m); }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }
inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }
inline uint32_t pool_size(uint32_t n, uint32_t k) {
    const uint64_t p = (uint64_t)k * 64u;
    uint32_t pool = p < n ? (uint32_t)p : n;
    return pool ? pool : 1u;
}
// Reserved for the radix sort's temporary storage: the size query needs a
// device, so workspace sizes use this bound and the call checks the real one.
inline size_t sort_temp_bound(uint32_t n) { return (size_t)n * 16 + (64ull << 20); }

struct KmLayout {
    size_t skeys, members, iota, sort_tmp, sort_tmp_bytes, sums, newc, offsets, maxbits, mean,
        pool_ix, closest, chunk_sum, u, empty_ids, draws, total;
};

KmLayout km_layout(uint32_t n, uint32_t dim, uint32_t k) {
    KmLayout L{};
    size_t off = 0;
    auto take = [&](size_t b) { const size_t o = off; off += align256(b); return o; };
    const uint32_t pool = pool_size(n, k);
    L.skeys = take((size_t)n * 4);
    L.members = take((size_t)n * 4);
    L.iota = take((size_t)n * 4);
    L.sort_tmp_bytes = sort_temp_bound(n);
    L.sort_tmp = take(L.sort_tmp_bytes);
    L.sums = take((size_t)k * dim * 4);
    L.newc = take((size_t)k * dim * 4);
    L.offsets = take(((size_t)k + 1) * 4);
    L.maxbits = take(4);
    L.mean = take((size_t)dim * 4);
    L.pool_ix = take((size_t)pool * 4);
    L.closest = take((size_t)pool * 4);
    L.chunk_sum = take(((size_t)pool + kSeedChunk - 1) / kSeedChunk * 8);
    L.u = take((size_t)k * 8);
    L.empty_ids = take((size_t)k * 4);
    L.draws = take((size_t)k * (kEmptyDraws + 1) * 4);
    L.total = off;
    return L;
}

#define FV_TRY(x)                            \
    do {                                     \
        const hipError_t e_ = (x);           \
        if (e_ != hipSuccess) return e_;     \
    } while (0)

hipError_t assign_launch(const float* X, uint32_t n, uint32_t dim, uint32_t ld, const float* C,
                         uint32_t k, uint32_t* assign, uint8_t* codes, uint32_t code_stride,
                         hipStream_t s) {
    const unsigned g = blocks_for(n, kAR);
    if (dim <= 8)
        km_assign_kernel<8><<<g, 256, 0, s>>>(X, n, dim, ld, C, k, assign, codes, code_stride);
    else if (dim <= 16)
        km_assign_kernel<16><<<g, 256, 0, s>>>(X, n, dim, ld, C, k, assign, codes, code_stride);
    else
        km_assign_kernel<32><<<g, 256, 0, s>>>(X, n, dim, ld, C, k, assign, codes, code_stride);
    return hipGetLastError();
}

// Stable grouping of rows by cluster: members (rows ascending within each
// cluster) and offsets [k + 1].
hipError_t group_by_cluster(const uint32_t* assign, uint32_t n, uint32_t k, char* ws,
                            const KmLayout& L, uint32_t* members, uint32_t* offsets,
                            hipStream_t s) {
    uint32_t* skeys = reinterpret_cast<uint32_t*>(ws + L.skeys);
    uint32_t* iota = reinterpret_cast<uint32_t*>(ws + L.iota);
    km_iota_kernel<<<blocks_for(n, 256), 256, 0, s>>>(iota, n);
    int end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) < k) ++end_bit;
    size_t tmp = 0;
    FV_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, assign, skeys, iota, members, n, 0,
                                              end_bit, s));
    if (tmp > L.sort_tmp_bytes) return hipErrorInvalidValue;
    FV_TRY(hipcub::DeviceRadixSort::SortPairs(ws + L.sort_tmp, tmp, assign, skeys, iota, members,
                                              n, 0, end_bit, s));
    km_offsets_kernel<<<blocks_for((size_t)n + 1, 256), 256, 0, s>>>(skeys, n, k, offsets);
    return hipGetLastError();
}

// `timed`: bracket the assign launches with the kernel timer (the dominant
// kernel of a coarse k-means).
hipError_t kmeans_fit(const float* X, uint32_t n, uint32_t dim, uint32_t ld, uint32_t k,
                      uint32_t max_iters, uint64_t seed, float* cent, uint32_t* assign,
                      uint32_t* sizes, uint32_t* offsets_out, uint32_t* members_out,
                      uint32_t* iters_run, char* ws, hipStream_t s, bool timed) {
    const KmLayout L = km_layout(n, dim, k);
    Rng rng{seed};
    // ---- k-means++ seeding over the pool (cluster.rs:46-93) ----
    const uint32_t pool = pool_size(n, k);
    std::vector<uint32_t> ix(n);
    for (uint32_t i = 0; i < n; ++i) ix[i] = i;
    for (uint32_t i = 0; i < pool; ++i) {
        const uint32_t j = i + rng.below(n - i);
        std::swap(ix[i], ix[j]);
    }
    std::vector<double> u(k, 0.0);
    for (uint32_t c = 1; c < k; ++c) u[c] = rng.unit();
    uint32_t* pool_ix = reinterpret_cast<uint32_t*>(ws + L.pool_ix);
    float* closest = reinterpret_cast<float*>(ws + L.closest);
    double* chunk_sum = reinterpret_cast<double*>(ws + L.chunk_sum);
    double* d_u = reinterpret_cast<double*>(ws + L.u);
    FV_TRY(hipMemcpyAsync(pool_ix, ix.data(), (size_t)pool * 4, hipMemcpyHostToDevice, s));
    FV_TRY(hipMemcpyAsync(d_u, u.data(), (size_t)k * 8, hipMemcpyHostToDevice, s));
    km_fill_f32_kernel<<<blocks_for(pool, 256), 256, 0, s>>>(closest, pool, __builtin_inff());
    km_copy_row_kernel<<<1, 64, 0, s>>>(X, ld, pool_ix, dim, cent);
    const uint32_t nch = (pool + kSeedChunk - 1) / kSeedChunk;
    for (uint32_t c = 1; c < k; ++c) {
        km_seed_update_kernel<<<nch, kSeedChunk, 0, s>>>(X, dim, ld, pool_ix, pool,
                                                         cent + (size_t)(c - 1) * dim, closest,
                                                         chunk_sum);
        km_seed_pick_kernel<<<1, 256, 0, s>>>(chunk_sum, nch, closest, pool, pool_ix, d_u + c, X,
                                             dim, ld, cent + (size_t)c * dim);
    }
    FV_TRY(hipGetLastError());

    // ---- Lloyd iterations (cluster.rs:95-180) ----
    uint32_t* members = members_out ? members_out : reinterpret_cast<uint32_t*>(ws + L.members);
    uint32_t* offsets = offsets_out ? offsets_out : reinterpret_cast<uint32_t*>(ws + L.offsets);
    float* sums = reinterpret_cast<float*>(ws + L.sums);
    float* newc = reinterpret_cast<float*>(ws + L.newc);
    uint32_t* maxbits = reinterpret_cast<uint32_t*>(ws + L.maxbits);
    float* mean = reinterpret_cast<float*>(ws + L.mean);
    fill_u32(assign, n, 0u, s);
    fill_u32(sizes, k, 0u, s);
    std::vector<uint32_t> h_sizes(k);
    bool have_mean = false;
    uint32_t done = 0;
    for (uint32_t it = 0; it < max_iters; ++it) {
        void* timer = timed ? timer_begin(s) : nullptr;
        FV_TRY(assign_launch(X, n, dim, ld, cent, k, assign, nullptr, 0, s));
        if (timer) timer_end(timer, s);
        FV_TRY(group_by_cluster(assign, n, k, ws, L, members, offsets, s));
        km_update_kernel<<<dim3(k, (dim + 63) / 64), 64, 0, s>>>(X, dim, ld, members, offsets,
                                                                 sums);
        fill_u32(maxbits, 1, 0u, s);
        km_finalize_kernel<<<blocks_for(k, 256), 256, 0, s>>>(sums, cent, offsets, k, dim, newc,
                                                              sizes, maxbits);
        uint32_t h_max = 0;
        FV_TRY(hipMemcpyAsync(h_sizes.data(), sizes, (size_t)k * 4, hipMemcpyDeviceToHost, s));
        FV_TRY(hipMemcpyAsync(&h_max, maxbits, 4, hipMemcpyDeviceToHost, s));
        FV_TRY(hipStreamSynchronize(s));
        std::vector<uint32_t> empties;
        for (uint32_t c = 0; c < k; ++c)
            if (h_sizes[c] == 0) empties.push_back(c);
        if (!empties.empty()) {
            if (!have_mean) {
                km_mean_kernel<<<blocks_for(dim, 64), 64, 0, s>>>(X, n, dim, ld, mean);
                have_mean = true;
            }
            const uint32_t ndraw = n < kEmptyDraws ? n : kEmptyDraws;
            std::vector<uint32_t> draws(empties.size() * (ndraw + 1));
            for (size_t e = 0; e < empties.size(); ++e)
                for (uint32_t t = 0; t <= ndraw; ++t) draws[e * (ndraw + 1) + t] = rng.below(n);
            uint32_t* d_ids = reinterpret_cast<uint32_t*>(ws + L.empty_ids);
            uint32_t* d_draws = reinterpret_cast<uint32_t*>(ws + L.draws);
            FV_TRY(hipMemcpyAsync(d_ids, empties.data(), empties.size() * 4,
                                  hipMemcpyHostToDevice, s));
            FV_TRY(hipMemcpyAsync(d_draws, draws.data(), draws.size() * 4, hipMemcpyHostToDevice,
                                  s));
            km_empty_kernel<<<(unsigned)empties.size(), 256, 0, s>>>(X, dim, ld, mean, d_ids,
                                                                     d_draws, ndraw, newc);
            // The host vectors above must outlive the copies.
            FV_TRY(hipStreamSynchronize(s));
        }
        FV_TRY(hipMemcpyAsync(cent, newc, (size_t)k * dim * 4, hipMemcpyDeviceToDevice, s));
        done = it + 1;
        float max_delta;
        memcpy(&max_delta, &h_max, 4);
        if (max_delta < 1e-4f) break;
    }
    if (done == 0) FV_TRY(group_by_cluster(assign, n, k, ws, L, members, offsets, s));
    FV_TRY(hipGetLastError());
    FV_TRY(hipStreamSynchronize(s));
    if (iters_run) *iters_run = done;
    return hipSuccess;
}

// The sub-space k-means runs are independent: up to kPqWorkers run at once,
// each from its own host thread on its own stream and workspace slot.
constexpr uint32_t kPqWorkers = 4;

struct BuildLayout {
    size_t km, residual, assign, sizes, pq_slot, pq_slot_bytes, pq_km, pq_assign, pq_sizes, total;
    uint32_t workers;
};

BuildLayout build_layout(uint32_t n, uint32_t dim, uint32_t nlist, uint32_t m_pq) {
    BuildLayout B{};
    const uint32_t nl = nlist > n ? n : (nlist ? nlist : 1u);
    const uint32_t kpq = n < 256u ? n : 256u;
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off += align256(bytes); return o; };
    B.km = take(km_layout(n, dim, nl).total);
    B.residual = take((size_t)n * dim * 4);
    B.assign = take((size_t)n * 4);
    B.sizes = take((size_t)nl * 4);
    B.workers = m_pq < kPqWorkers ? m_pq : kPqWorkers;
    size_t so = 0;
    auto slot = [&](size_t bytes) { const size_t o = so; so += align256(bytes); return o; };
    B.pq_km = slot(km_layout(n, dim / m_pq, kpq).total);
    B.pq_assign = slot((size_t)n * 4);
    B.pq_sizes = slot(256 * 4);
    B.pq_slot_bytes = so;
    B.pq_slot = take(so * B.workers);
    B.total = off;
    return B;
}

}  // namespace
}  // namespace fv

using namespace fv;

extern "C" {

int fv_kmeans_assign(const float* d_x, uint32_t n, uint32_t dim, uint32_t ld,
                     const float* d_centroids, uint32_t k, uint32_t* d_assign, uint8_t* d_codes,
                     uint32_t code_stride, void* stream) {
    if (!d_x || !d_centroids || (!d_assign && !d_codes)) return -1;
    if (n == 0) return 0;
    if (dim == 0 || k == 0 || ld < dim || (d_codes && k > 256u)) return (int)hipErrorInvalidValue;
    return (int)assign_launch(d_x, n, dim, ld, d_centroids, k, d_assign, d_codes, code_stride,
                              (hipStream_t)stream);
}

size_t fv_kmeans_workspace_size(uint32_t n, uint32_t dim, uint32_t k) {
    if (n == 0 || dim == 0 || k == 0 || k > n) return 0;
    return km_layout(n, dim, k).total;
}

int fv_kmeans_fit(const float* d_x, uint32_t n, uint32_t dim, uint32_t ld, uint32_t k,
                  uint32_t max_iters, uint64_t seed, float* d_centroids, uint32_t* d_assign,
                  uint32_t* d_sizes, uint32_t* d_offsets, uint32_t* d_members,
                  uint32_t* iters_run, void* d_workspace, size_t workspace_bytes, void* stream) {
    if (!d_x || !d_centroids || !d_assign || !d_sizes || !d_workspace) return -1;
    if (n == 0 || dim == 0 || k == 0 || k > n || ld < dim) return (int)hipErrorInvalidValue;
    if (workspace_bytes < km_layout(n, dim, k).total) return (int)hipErrorInvalidValue;
    return (int)kmeans_fit(d_x, n, dim, ld, k, max_iters, seed, d_centroids, d_assign, d_sizes,
                           d_offsets, d_members, iters_run, static_cast<char*>(d_workspace),
                           (hipStream_t)stream, true);
}

size_t fv_ivf_pq_build_workspace_size(uint32_t n, uint32_t dim, uint32_t nlist, uint32_t m_pq) {
    if (n == 0 || dim == 0 || m_pq == 0 || dim % m_pq) return 0;
    return build_layout(n, dim, nlist, m_pq).total;
}

int fv_ivf_pq_build(const float* d_vectors, uint32_t n, uint32_t dim, uint32_t nlist,
                    uint32_t m_pq, uint32_t coarse_iters, uint32_t pq_iters, uint64_t seed,
                    float* d_centroids, float* d_codebooks, uint8_t* d_codes,
                    uint32_t* d_ivf_offsets, uint32_t* d_ivf_vec_ids, uint32_t* iters_run,
                    void* d_workspace, size_t workspace_bytes, void* stream) {
    if (!d_vectors || !d_centroids || !d_codebooks || !d_codes || !d_ivf_offsets ||
        !d_ivf_vec_ids || !d_workspace)
        return -1;
    if (n == 0 || dim == 0 || m_pq == 0 || dim % m_pq) return (int)hipErrorInvalidValue;
    const BuildLayout B = build_layout(n, dim, nlist, m_pq);
    if (workspace_bytes < B.total) return (int)hipErrorInvalidValue;
    hipStream_t s = (hipStream_t)stream;
    char* ws = static_cast<char*>(d_workspace);
    const uint32_t nl = nlist > n ? n : (nlist ? nlist : 1u);
    const uint32_t sub = dim / m_pq;
    const uint32_t kpq = n < 256u ? n : 256u;
    uint32_t* assign = reinterpret_cast<uint32_t*>(ws + B.assign);
    uint32_t* sizes = reinterpret_cast<uint32_t*>(ws + B.sizes);
    float* res = reinterpret_cast<float*>(ws + B.residual);
    hipError_t e = kmeans_fit(d_vectors, n, dim, dim, nl, coarse_iters, seed, d_centroids, assign,
                              sizes, d_ivf_offsets, d_ivf_vec_ids, iters_run, ws + B.km, s, true);
    if (e != hipSuccess) return (int)e;
    const size_t total = (size_t)n * dim;
    km_residual_kernel<<<blocks_for(total, 256), 256, 0, s>>>(d_vectors, d_centroids, assign, n,
                                                              dim, res);
    e = hipStreamSynchronize(s);  // residuals ready for the worker streams
    if (e != hipSuccess) return (int)e;
    int device = 0;
    e = hipGetDevice(&device);
    if (e != hipSuccess) return (int)e;
    std::vector<hipError_t> errs(B.workers, hipSuccess);
    auto worker = [&](uint32_t w) {
        hipError_t err = hipSetDevice(device);
        hipStream_t ws_stream = nullptr;
        if (err == hipSuccess) err = hipStreamCreateWithFlags(&ws_stream, hipStreamNonBlocking);
        char* slot = ws + B.pq_slot + (size_t)w * B.pq_slot_bytes;
        uint32_t* sub_assign = reinterpret_cast<uint32_t*>(slot + B.pq_assign);
        uint32_t* sub_sizes = reinterpret_cast<uint32_t*>(slot + B.pq_sizes);
        for (uint32_t m = w; err == hipSuccess && m < m_pq; m += B.workers) {
            float* cb = d_codebooks + (size_t)m * 256 * sub;
            err = kmeans_fit(res + (size_t)m * sub, n, sub, dim, kpq, pq_iters,
                             seed + (uint64_t)(m + 1) * 0x9E3779B97F4A7C15ull, cb, sub_assign,
                             sub_sizes, nullptr, nullptr, iters_run ? iters_run + 1 + m : nullptr,
                             slot + B.pq_km, ws_stream, false);
            if (err != hipSuccess) break;
            if (kpq < 256u)
                km_fill_f32_kernel<<<blocks_for((size_t)(256 - kpq) * sub, 256), 256, 0,
                                     ws_stream>>>(cb + (size_t)kpq * sub,
                                                  (size_t)(256 - kpq) * sub, FLT_MAX);
            err = assign_launch(res + (size_t)m * sub, n, sub, dim, cb, 256, nullptr,
                                d_codes + m, m_pq, ws_stream);
        }
        if (ws_stream) {
            const hipError_t se = hipStreamSynchronize(ws_stream);
            if (err == hipSuccess) err = se;
            (void)hipStreamDestroy(ws_stream);
        }
        errs[w] = err;
    };
    std::vector<std::thread> pool;
    uint32_t started = 1;
    for (uint32_t w = 1; w < B.workers; ++w) {
        try {
            pool.emplace_back(worker, w);
        } catch (const std::system_error&) {
            break;  // no more threads: the remaining slots run on this one
        }
        started = w + 1;
    }
    worker(0);
    for (uint32_t w = started; w < B.workers; ++w) worker(w);
    for (auto& t : pool) t.join();
    for (hipError_t we : errs)
        if (we != hipSuccess) return (int)we;
    return 0;
}

}  // extern "C"
