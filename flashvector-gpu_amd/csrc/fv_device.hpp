// Device-side building blocks shared by the scan kernels (gfx950, wave64).
//
// Results are ordered everywhere by the 64-bit key  ord(distance) << 32 | slot,
// where ord() is a monotone map of IEEE binary32 onto uint32. Sorting keys
// therefore sorts by (distance, slot) with ties going to the lower slot, the
// rule of the reference CPU path (crates/engine/src/metrics.rs:125-129) and of
// its flat scan kernel's strict-< insertion (kernels/src/sparse_hybrid.cu:60-73).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/flashvector/types.h"

namespace fv {

constexpr int kWave = 64;
constexpr uint64_t kEmptyKey = ~0ull;
constexpr float kAdmit = 1e30f;  // only distances below the reference's sentinel are kept

__device__ __forceinline__ uint32_t ord_f32(float f) {
    f = f + 0.0f;  // -0 -> +0 so that both zeros share one key
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float unord_f32(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(u);
}

__device__ __forceinline__ uint64_t make_key(float dist, uint32_t slot) {
    return ((uint64_t)ord_f32(dist) << 32) | slot;
}

__device__ __forceinline__ float key_dist(uint64_t k) { return unord_f32((uint32_t)(k >> 32)); }
__device__ __forceinline__ uint32_t key_slot(uint64_t k) { return (uint32_t)k; }

// Buffer resource over `bytes` bytes at p (raw, 32-bit elements): loads past
// the end return zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                             0x00020000);
}

__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// kernels/src/distance_metrics.cuh:17-20: LSB-first u32 words, NULL allows all.
__device__ __forceinline__ bool mask_allows(const uint32_t* __restrict__ mask, uint32_t idx) {
    return mask == nullptr || ((__ldg(mask + (idx >> 5)) >> (idx & 31u)) & 1u) != 0u;
}

__host__ __device__ __forceinline__ uint32_t next_pow2(uint32_t v) {
    if (v <= 1u) return 1u;
    --v;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1u;
}

// Orders LDS operations of one wavefront: lanes of a wave run in lockstep and
// the LDS unit serves one wave's requests in order, so a compiler fence plus a
// wave barrier is enough between dependent bitonic stages.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ascending bitonic sort of keys[0..n) (n a power of two) by one wavefront.
__device__ __forceinline__ void wave_bitonic_sort(uint64_t* keys, uint32_t n, int lane) {
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t p = (uint32_t)lane; p < (n >> 1); p += kWave) {
                uint32_t i = ((p & ~(j - 1u)) << 1) | (p & (j - 1u));
                uint32_t m = i + j;
                uint64_t a = keys[i], b = keys[m];
                bool up = (i & k) == 0;
                if ((a > b) == up) {
                    keys[i] = b;
                    keys[m] = a;
                }
            }
            wave_lds_sync();
        }
    }
}

// Same network executed by a whole workgroup (blockDim.x threads).
__device__ __forceinline__ void block_bitonic_sort(uint64_t* keys, uint32_t n) {
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t p = threadIdx.x; p < (n >> 1); p += blockDim.x) {
                uint32_t i = ((p & ~(j - 1u)) << 1) | (p & (j - 1u));
                uint32_t m = i + j;
                uint64_t a = keys[i], b = keys[m];
                bool up = (i & k) == 0;
                if ((a > b) == up) {
                    keys[i] = b;
                    keys[m] = a;
                }
            }
            __syncthreads();
        }
    }
}

// Same network on (key, value) pairs ordered lexicographically.
__device__ __forceinline__ void block_bitonic_sort_pairs(uint64_t* keys, uint32_t* vals, uint32_t n) {
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t p = threadIdx.x; p < (n >> 1); p += blockDim.x) {
                uint32_t i = ((p & ~(j - 1u)) << 1) | (p & (j - 1u));
                uint32_t m = i + j;
                const uint64_t a = keys[i], b = keys[m];
                const uint32_t va = vals[i], vb = vals[m];
                const bool gt = a > b || (a == b && va > vb);
                bool up = (i & k) == 0;
                if (gt == up) {
                    keys[i] = b;
                    keys[m] = a;
                    vals[i] = vb;
                    vals[m] = va;
                }
            }
            __syncthreads();
        }
    }
}

// Same network on (key, value) pairs by one wavefront.
__device__ __forceinline__ void wave_bitonic_sort_pairs(uint64_t* keys, uint32_t* vals, uint32_t n,
                                                        int lane) {
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t p = (uint32_t)lane; p < (n >> 1); p += kWave) {
                uint32_t i = ((p & ~(j - 1u)) << 1) | (p & (j - 1u));
                uint32_t m = i + j;
                const uint64_t a = keys[i], b = keys[m];
                const uint32_t va = vals[i], vb = vals[m];
                const bool gt = a > b || (a == b && va > vb);
                bool up = (i & k) == 0;
                if (gt == up) {
                    keys[i] = b;
                    keys[m] = a;
                    vals[i] = vb;
                    vals[m] = va;
                }
            }
            wave_lds_sync();
        }
    }
}

// Inclusive prefix sum of v over the lanes of a wavefront.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t t = __shfl_up(v, off);
        if (lane >= (uint32_t)off) v += t;
    }
    return v;
}

// ---- Workgroup radix selection ---------------------------------------------
// Workgroups of 256 to 1024 threads (whole wavefronts).
constexpr uint32_t kSelThreads = 256;    // the IVF scan's workgroup
constexpr uint32_t kSelMaxWaves = 16;

// LDS scratch of the selection helpers below (one per workgroup).
struct BlockSelectLds {
    uint64_t thr;                     // key stored by block_compact's keep == 2
    uint64_t red[2 * kSelMaxWaves];   // per-wave smallest / largest key
    uint32_t sel[3];                  // radix round: digit, keys below it, keys in it
    uint32_t wsum[2][kSelMaxWaves];   // per-wave survivor counts
    uint32_t hist[2][256];            // digit histograms of consecutive rounds
};

// Radix selection over the c distinct keys buf[0..c) (c > keep >= 1): the
// bits above the highest bit in which the smallest and largest key differ
// are common to all keys; below it, 8-bit digits are histogrammed one round
// at a time over the keys that share the prefix found so far, until the
// digit holding the keep-th smallest holds no other key (or every bit is
// fixed). On return exactly `keep` keys satisfy (key & pmask) <= prefix, and
// the keep-th smallest is the one with (key & pmask) == prefix. All threads
// call it after a barrier; buf is only read.
__device__ inline void block_radix_kth(const uint64_t* buf, uint32_t c, uint32_t keep,
                                       BlockSelectLds& L, uint64_t& pmask_out,
                                       uint64_t& prefix_out) {
    const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const uint32_t nt = blockDim.x, nw = nt / kWave;
    unsigned long long lo = ~0ull, hi = 0;
    for (uint32_t i = tid; i < c; i += nt) {
        const unsigned long long k = buf[i];
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const unsigned long long a = __shfl_xor(lo, off), b = __shfl_xor(hi, off);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if (lane == 0) {
        L.red[wave] = lo;
        L.red[kSelMaxWaves + wave] = hi;
    }
    for (uint32_t t = tid; t < 256u; t += nt) L.hist[0][t] = 0;
    __syncthreads();
    lo = L.red[0];
    hi = L.red[kSelMaxWaves];
    for (uint32_t w = 1; w < nw; ++w) {
        lo = L.red[w] < lo ? L.red[w] : lo;
        hi = L.red[kSelMaxWaves + w] > hi ? L.red[kSelMaxWaves + w] : hi;
    }
    uint32_t hb = 64u - (uint32_t)__clzll((long long)(lo ^ hi));  // undetermined low bits
    uint64_t pmask = hb >= 64u ? 0ull : ~0ull << hb;
    uint64_t prefix = lo & pmask;
    uint32_t need = keep;  // rank of the wanted key among those sharing the prefix
    for (uint32_t r = 0;; ++r) {
        const uint32_t nb = hb < 8u ? hb : 8u, shift = hb - nb;
        const uint64_t dmask = (1ull << nb) - 1ull;
        uint32_t* h = L.hist[r & 1u];
        for (uint32_t i = tid; i < c; i += nt) {
            const uint64_t k = buf[i];
            if ((k & pmask) == prefix) atomicAdd(&h[(uint32_t)((k >> shift) & dmask)], 1u);
        }
        for (uint32_t t = tid; t < 256u; t += nt) L.hist[(r + 1u) & 1u][t] = 0;
        __syncthreads();
        if (wave == 0) {
            const uint32_t h0 = h[4 * lane], h1 = h[4 * lane + 1], h2 = h[4 * lane + 2],
                           h3 = h[4 * lane + 3];
            const uint32_t sum = h0 + h1 + h2 + h3;
            const uint32_t incl = wave_incl_scan(sum, lane), excl = incl - sum;
            if (excl < need && need <= incl) {
                uint32_t b = 4 * lane, before = excl, cnt = h0;
                if (before + h0 < need) {
                    before += h0; ++b; cnt = h1;
                    if (before + h1 < need) {
                        before += h1; ++b; cnt = h2;
                        if (before + h2 < need) {
                            before += h2; ++b; cnt = h3;
                        }
                    }
                }
                L.sel[0] = b;
                L.sel[1] = before;
                L.sel[2] = cnt;
            }
        }
        __syncthreads();
        const uint32_t b = L.sel[0], cnt = L.sel[2];
        need -= L.sel[1];
        prefix |= (uint64_t)b << shift;
        pmask |= dmask << shift;
        hb = shift;
        if (cnt == 1u || hb == 0u) break;
    }
    pmask_out = pmask;
    prefix_out = prefix;
}

// In-place compaction of buf[0..c) to the keys for which keepf(key) is
// nonzero, in no particular order; a return of 2 also stores the key in
// L.thr. Returns the number kept (the same in every thread). 8 keys per
// thread per chunk: every key of a chunk is read before the chunk's barrier
// and survivors are written after it, to positions below the number of keys
// read so far. All threads call it after a barrier; callers barrier before
// other threads read the survivors.
template <class F>
__device__ inline uint32_t block_compact(uint64_t* buf, uint32_t c, F keepf, BlockSelectLds& L) {
    const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const uint32_t nt = blockDim.x, nw = nt / kWave;
    uint32_t running = 0;
    for (uint32_t base = 0, ch = 0; base < c; base += 8u * nt, ++ch) {
        uint64_t v[8];
        uint32_t keepmask = 0, n = 0;
#pragma unroll
        for (uint32_t u = 0; u < 8u; ++u) {
            const uint32_t i = base + u * nt + tid;
            v[u] = i < c ? buf[i] : ~0ull;
            if (i < c) {
                const int f = keepf(v[u]);
                if (f) {
                    keepmask |= 1u << u;
                    ++n;
                    if (f == 2) L.thr = v[u];
                }
            }
        }
        const uint32_t incl = wave_incl_scan(n, lane);
        if (lane == kWave - 1) L.wsum[ch & 1u][wave] = incl;
        __syncthreads();
        uint32_t off = running + incl - n;
        for (uint32_t w = 0; w < nw; ++w) {
            const uint32_t ws = L.wsum[ch & 1u][w];
            off += w < wave ? ws : 0u;
            running += ws;
        }
#pragma unroll
        for (uint32_t u = 0; u < 8u; ++u)
            if (keepmask & (1u << u)) buf[off++] = v[u];
    }
    return running;
}

// Keeps the `keep` smallest of the c distinct keys buf[0..c) (c > keep >= 1),
// compacted into buf[0..keep) in no particular order, and returns the keep-th
// smallest to every thread (block_radix_kth, then block_compact). Same call
// and barrier rules as block_compact.
__device__ inline uint64_t block_select(uint64_t* buf, uint32_t c, uint32_t keep,
                                        BlockSelectLds& L) {
    uint64_t pmask, prefix;
    block_radix_kth(buf, c, keep, L, pmask, prefix);
    block_compact(buf, c, [pmask, prefix](uint64_t k) {
        const uint64_t mp = k & pmask;
        return mp < prefix ? 1 : (mp == prefix ? 2 : 0);
    }, L);
    return L.thr;
}

// Prune of one query's candidate buffer by one wavefront: sorts the `count`
// buffered keys through LDS scratch (capacity >= next_pow2(count)), writes the
// `keep` smallest back in ascending order and returns how many remain.
// `scratch` is wave-private LDS.
__device__ __forceinline__ uint32_t wave_prune(uint64_t* __restrict__ buf, uint32_t count,
                                               uint32_t keep, uint64_t* scratch, int lane) {
    uint32_t n = next_pow2(count < 2u ? 2u : count);
    for (uint32_t i = (uint32_t)lane; i < n; i += kWave) scratch[i] = i < count ? buf[i] : kEmptyKey;
    wave_lds_sync();
    wave_bitonic_sort(scratch, n, lane);
    uint32_t kept = count < keep ? count : keep;
    for (uint32_t i = (uint32_t)lane; i < kept; i += kWave) buf[i] = scratch[i];
    wave_lds_sync();
    return kept;
}

// Exact prune of one query's candidate buffer (count <= 64*R distinct keys,
// global memory) by one wavefront in registers: bisection over the whole
// 64-bit key finds the keep-th smallest key t (keys are distinct, so exactly
// `keep` keys are at or below it, however many share its distance), then the
// keys above t are dropped by an in-order compaction. Returns the new count;
// *bound = the distance bits of t (0xFFFFFFFF when nothing was dropped).
template <int R>
__device__ __forceinline__ uint32_t wave_key_prune(uint64_t* __restrict__ buf, uint32_t count,
                                                   uint32_t keep, int lane, uint32_t* bound) {
    uint64_t v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = (uint32_t)(r * kWave + lane);
        v[r] = i < count ? buf[i] : kEmptyKey;
    }
    if (count <= keep) {
        *bound = 0xFFFFFFFFu;
        return count;
    }
    uint64_t lo = 0u, hi = kEmptyKey - 1u;  // buffered keys are below kEmptyKey
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        uint32_t c = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) c += __popcll(__ballot(v[r] <= mid));
        if (c >= keep) hi = mid;
        else lo = mid + 1u;
    }
    uint32_t base = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool k = v[r] <= hi;
        const uint64_t m = __ballot(k);
        const uint32_t pos = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (k) buf[pos] = v[r];
        base += __popcll(m);
    }
    *bound = (uint32_t)(hi >> 32);
    return base;
}

// Prune of one query's candidate buffer (count <= 64*R keys in global memory)
// by one wavefront without sorting: bisection over the ordered distance
// finds a bound t with at least `keep` keys at or below it (ballot +
// popcount per probe), then the keys above t are dropped by an in-order
// compaction (mbcnt prefix). Every dropped key has at least `keep` kept keys
// at or below its distance, and all of them come from earlier rows, so none
// of the dropped keys (nor any later row at distance >= t) can enter the
// chunk's top-`keep`. Returns the new count; *bound = t.
template <int R>
__device__ __forceinline__ uint32_t wave_bisect_prune(uint64_t* __restrict__ buf, uint32_t count,
                                                      uint32_t keep, int lane, uint32_t* bound) {
    uint64_t v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = (uint32_t)(r * kWave + lane);
        v[r] = i < count ? buf[i] : kEmptyKey;
    }
    if (count <= keep) {
        *bound = 0xFFFFFFFFu;
        return count;
    }
    // Smallest t with |{key : ord(key) <= t}| >= keep, stopping early once the
    // kept set is within one wavefront of keep.
    uint32_t lo = 0u, hi = 0xFFFFFFFEu;
    uint32_t t = hi;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        uint32_t c = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) c += __popcll(__ballot((uint32_t)(v[r] >> 32) <= mid));
        if (c >= keep) {
            hi = mid;
            if (c <= keep + kWave) break;
        } else {
            lo = mid + 1u;
        }
    }
    t = hi;
    // In-order compaction of the keys at or below t.
    uint32_t base = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool k = (uint32_t)(v[r] >> 32) <= t;
        const uint64_t m = __ballot(k);
        const uint32_t pos = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (k) buf[pos] = v[r];
        base += __popcll(m);
    }
    *bound = t;
    return base;
}

// Exact FP32 distances in the reference CPU path's order: one sequential,
// non-fused sum over d (the library is built with -ffp-contract=off).
template <int METRIC>
__device__ __forceinline__ float exact_distance(const float* __restrict__ q,
                                                const float* __restrict__ x, uint32_t dim) {
    if (METRIC == METRIC_INNER_PRODUCT) {
        float dot = 0.0f;
        for (uint32_t d = 0; d < dim; ++d) {
            float p = q[d] * x[d];
            dot = dot + p;
        }
        return -dot;
    } else if (METRIC == METRIC_COSINE) {
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
    } else {
        float acc = 0.0f;
        for (uint32_t d = 0; d < dim; ++d) {
            float diff = q[d] - x[d];
            float sq = diff * diff;
            acc = acc + sq;
        }
        return acc;
    }
}

// Same sums with 128-bit row loads (the order over d is unchanged); x must be
// 16-byte aligned and dim a multiple of 4, else the scalar form runs.
template <int METRIC>
__device__ __forceinline__ float exact_distance_vec(const float* __restrict__ q,
                                                    const float* __restrict__ x, uint32_t dim) {
    if ((dim & 3u) || (reinterpret_cast<uintptr_t>(x) & 15u) || (reinterpret_cast<uintptr_t>(q) & 15u))
        return exact_distance<METRIC>(q, x, dim);
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const float4* q4 = reinterpret_cast<const float4*>(q);
    float a = 0.0f, nq = 0.0f, nx = 0.0f;
#pragma unroll 4
    for (uint32_t i = 0; i < dim / 4; ++i) {
        const float4 xv = x4[i], qv = q4[i];
        const float xs[4] = {xv.x, xv.y, xv.z, xv.w};
        const float qs[4] = {qv.x, qv.y, qv.z, qv.w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (METRIC == METRIC_L2) {
                float diff = qs[c] - xs[c];
                float sq = diff * diff;
                a = a + sq;
            } else {
                float p = qs[c] * xs[c];
                a = a + p;
                if (METRIC == METRIC_COSINE) {
                    float pa = qs[c] * qs[c], pb = xs[c] * xs[c];
                    nq = nq + pa;
                    nx = nx + pb;
                }
            }
        }
    }
    if (METRIC == METRIC_L2) return a;
    if (METRIC == METRIC_INNER_PRODUCT) return -a;
    float denom = sqrtf(nq) * sqrtf(nx) + 1e-8f;
    return 1.0f - a / denom;
}

__device__ __forceinline__ float exact_distance_rt(const float* q, const float* x, uint32_t dim,
                                                   int metric) {
    if (metric == METRIC_INNER_PRODUCT) return exact_distance<METRIC_INNER_PRODUCT>(q, x, dim);
    if (metric == METRIC_COSINE) return exact_distance<METRIC_COSINE>(q, x, dim);
    return exact_distance<METRIC_L2>(q, x, dim);
}

// The same distances evaluated in FP64 (products of floats are exact there).
template <int METRIC>
__device__ __forceinline__ double exact_distance_f64(const float* __restrict__ q,
                                                     const float* __restrict__ x, uint32_t dim) {
    double a = 0.0, nq = 0.0, nx = 0.0;
    for (uint32_t d = 0; d < dim; ++d) {
        const double qd = q[d], xd = x[d];
        if (METRIC == METRIC_L2) {
            const double t = qd - xd;
            a = __builtin_fma(t, t, a);
        } else {
            a = __builtin_fma(qd, xd, a);
            if (METRIC == METRIC_COSINE) {
                nq = __builtin_fma(qd, qd, nq);
                nx = __builtin_fma(xd, xd, nx);
            }
        }
    }
    if (METRIC == METRIC_L2) return a;
    if (METRIC == METRIC_INNER_PRODUCT) return -a;
    return 1.0 - a / (sqrt(nq) * sqrt(nx) + 1e-8);
}

// OCP FP8 E4M3 (bias 7, no infinities, largest finite 448): round to nearest
// even, saturating to +-448; NaN maps to the NaN code 0x7f.
// FP16 bits of a stored row element, by the reference's host encoder
// (crates/engine/src/quant.rs:237-273): normal results round half away from
// zero on the magnitude (mantissa + 0x1000, then truncate), subnormal
// results truncate, overflow gives infinity, NaN becomes the quiet 0x7e00.
// Stored codes must equal the reference's tables bit for bit; queries are
// converted with round-to-nearest-even, as the reference's kernel does.
__host__ __device__ __forceinline__ uint16_t f32_to_f16_rows(float x) {
    uint32_t bits;
    __builtin_memcpy(&bits, &x, 4);
    const uint16_t sign = (uint16_t)((bits >> 16) & 0x8000u);
    const int32_t exp = (int32_t)((bits >> 23) & 0xffu);
    const uint32_t mant = bits & 0x007fffffu;
    if (exp == 255) return (uint16_t)(sign | (mant == 0 ? 0x7c00u : 0x7e00u));
    const int32_t e16 = exp - 112;
    if (e16 >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e16 <= 0) {
        if (e16 < -10) return sign;
        return (uint16_t)(sign | (uint16_t)((mant | 0x00800000u) >> (14 - e16)));
    }
    const uint32_t r = mant + 0x1000u;
    if (r & 0x00800000u) {
        return e16 + 1 >= 31 ? (uint16_t)(sign | 0x7c00u) : (uint16_t)(sign | (uint16_t)((e16 + 1) << 10));
    }
    return (uint16_t)(sign | (uint16_t)(e16 << 10) | (uint16_t)(r >> 13));
}

__host__ __device__ __forceinline__ uint8_t f32_to_e4m3(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    const uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    u &= 0x7fffffffu;
    if (u > 0x7f800000u) return (uint8_t)(sign | 0x7fu);
    const float a = __builtin_bit_cast(float, u);
    if (a >= 448.0f) return (uint8_t)(sign | 0x7eu);
    const int e = (int)(u >> 23) - 127;
    if (e < -6) {
        // Subnormal range (step 2^-9); m == 8 is the code of 2^-6.
        const uint32_t m = (uint32_t)rintf(a * 512.0f);
        return (uint8_t)(sign | m);
    }
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

}  // namespace fv
