// Matrix-core scan kernel of the reduced-precision flat search, shared by the
// per-dtype translation units (fv_flat_mfma_f16.hip, _bf16.hip, _fp8.hip),
// which instantiate it for their K-step counts.
//
// Reference: kernels/src/ivf_quant_lookup.cu:887-1058 (wmma 16x16x16, one
// block per 16 queries, FP32->FP16 query tiles reloaded per k-slice, serial
// insertion epilogue, no rerank).
//
// Design (gfx950, wave64):
//  * 16-bit codes: v_mfma_f32_16x16x32_{f16,bf16}, A = 16 database rows,
//    B = 16 negated queries, C = 0.5*|x|^2 of each row, so D = 0.5|x|^2 - q.x
//    and the L2 distance is |q|^2 + 2D: the |x|^2 term costs nothing.
//  * FP8 e4m3 codes with one FP32 scale per row and per query:
//    v_mfma_f32_16x16x128_f8f6f4 accumulates code products from zero and the
//    epilogue applies D = c_row + (s_row * s_query) * acc.
//  * Masked and out-of-range rows get c_row = +inf and never pass a threshold.
//  * A workgroup owns 4 waves x QW queries (QW = 64, or 32 or 16 when more
//    queries' fragments would not fit the register file) and one chunk of the row
//    space. The queries' B fragments live in registers for the whole kernel;
//    row tiles are double-buffered through LDS, one barrier per tile. The K
//    chunks of a step are interleaved over the four lane groups and the tile
//    rows padded or XOR-swizzled so that the fragment reads are free of bank
//    conflicts, register-staged FP8 tiles excepted (ds_read_b128 serves 16 lanes per cycle from 64 banks: lanes
//    {0-3, 12-15, 20-23, 24-27}, {4-11, 16-19, 28-31}, and the same pattern
//    in the upper half; measured with the LDS conflict counter).
//  * The row space is either every row (the mask, if any, is tested per row)
//    or a compacted ascending list of the valid rows (sparse masks). Chunk
//    bounds are computed on the device from the actual row count.
//  * Selection: survivors of the per-query threshold are compacted wave-wide
//    (ballot + mbcnt) into an LDS staging list and flushed into per-(chunk,
//    query) buffers in global memory. Near-full buffers are pruned to the
//    query's shortlist depth S, by the owning wave (register bisection) or,
//    for deep shortlists, by the workgroup. Thresholds combine the chunk's
//    own S-th best (strict: rows arrive in ascending order) with a per-query
//    bound shared by all chunks through global memory (non-strict: the S-th
//    best of any chunk bounds the global S-th best), prefetched every tile.
//  * Workgroup ids are mapped XCD-first so that the workgroups of one XCD
//    sweep the same chunk together and share it through that XCD's L2.
#pragma once

#include <type_traits>

#include "fv_flat_mfma.hpp"

namespace fv {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef int int32x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaThreads = 256;
// Staged candidate records per wave: a lane whose 4 scores for one query
// column have a minimum under the threshold writes them (16 B) and their
// position (4 B); keys are built when the records are flushed.
constexpr uint32_t kRec = 128;
constexpr uint32_t kRecBytes = kRec * 20;
// Workgroup-prune instantiations: the radix selection's LDS block, then the
// four waves' record staging (16-byte aligned).
constexpr uint32_t kScanSelBytes = (uint32_t)((sizeof(BlockSelectLds) + 15) / 16 * 16);

template <int DT>
struct MfmaOps;

template <>
struct MfmaOps<FV_DTYPE_F16> {
    using vec = half8_t;
    static constexpr int kLaneBytes = 16;  // 8 halves of K per lane per step
    static __device__ __forceinline__ floatx4 mma(vec a, vec b, floatx4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct MfmaOps<FV_DTYPE_BF16> {
    using vec = bf16x8_t;
    static constexpr int kLaneBytes = 16;
    static __device__ __forceinline__ floatx4 mma(vec a, vec b, floatx4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

template <>
struct MfmaOps<FV_DTYPE_FP8_E4M3> {
    using vec = int32x8_t;
    static constexpr int kLaneBytes = 32;  // 32 e4m3 codes of K per lane per step
    static __device__ __forceinline__ floatx4 mma(vec a, vec b, floatx4 c) {
        // Formats 0/0 = e4m3 for both operands; zero scale operands select
        // the unscaled instruction.
        return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, 0, 0, 0, 0);
    }
};

// Compile-time shape of one instantiation.
template <int DT, int KS>
struct ScanShape {
    static constexpr int kLaneBytes = MfmaOps<DT>::kLaneBytes;
    static constexpr int kStepBytes = 4 * kLaneBytes;  // bytes of one row per K step
    static constexpr int kRowBytes = KS * kStepBytes;
    // Queries per wave: 64, 32 or 16, the most whose fragments take at most
    // 192 VGPRs (16 queries reach 1536 dims of 16-bit codes, 3072 of FP8).
    static constexpr int kRegs16 = KS * (kLaneBytes / 4);  // VGPRs per 16 queries
    static constexpr int kQW = 4 * kRegs16 <= 192 ? 64 : (2 * kRegs16 <= 192 ? 32 : 16);
    static constexpr int kQB = 4 * kQW;
    static constexpr int kBRegs = (kQW / 16) * KS * (kLaneBytes / 4);
    static constexpr int kOcc = kBRegs < 128 ? 2 : 1;
    // Rows per LDS tile: 128, 64, 32 or 16, the most whose two tiles fit
    // 100 KB (taller tiles beat a second workgroup per CU: measured at 768
    // dims).
    static constexpr int kRT = 2 * 128 * (kRowBytes + 16) <= 102400
                                   ? 128
                                   : (2 * 64 * (kRowBytes + 16) <= 102400
                                          ? 64
                                          : (2 * 32 * (kRowBytes + 16) <= 102400 ? 32 : 16));
    // Workgroup-prune instantiations (deep shortlists) need up to 82 KB of
    // prune scratch, so their double-buffered tiles stay within 50 KB.
    static constexpr int kRTDeep = 2 * 64 * (kRowBytes + 16) <= 51200
                                       ? 64
                                       : (2 * 32 * (kRowBytes + 16) <= 51200 ? 32 : 16);
};

// Rows of whole 256-byte multiples can be staged with direct-to-LDS loads
// into an unpadded, XOR-swizzled tile (see flat_mfma_scan_kernel).
template <int DT, int KS>
constexpr bool scan_dma_ok() { return ScanShape<DT, KS>::kRowBytes % 256 == 0; }

// One 16-byte direct-to-LDS load per lane: lane l's 16 bytes at byte offset
// `off` of the buffer land at lds_base + 16 * l (lds_base uniform across the
// wavefront). A buffer (not global) load: the compiler then tracks it on the
// vector-memory counter only, so LDS reads issued after it keep counted
// (lgkmcnt(N)) waits instead of waiting for every outstanding LDS read.
__device__ __forceinline__ void lds_dma16(__amdgpu_buffer_rsrc_t rsrc, uint32_t off,
                                          void* lds_base) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(
        rsrc, (__attribute__((address_space(3))) void*)lds_base, 16, off, 0, 0, 0);
}

// s_waitcnt immediate of the gfx9 family (gfx950): vmcnt in bits 3:0 and
// 15:14, expcnt in bits 6:4, lgkmcnt in bits 11:8; a counter given its field's
// maximum is not waited on.
constexpr int waitcnt_imm(int vmcnt, int expcnt, int lgkmcnt) {
    return (vmcnt & 0xF) | ((vmcnt >> 4) << 14) | ((expcnt & 0x7) << 4) | ((lgkmcnt & 0xF) << 8);
}
constexpr int kWaitVmcnt0 = waitcnt_imm(0, 7, 15);  // every vector-memory load retired
static_assert(kWaitVmcnt0 == 0x0F70, "vmcnt(0) expcnt(7) lgkmcnt(15)");

// Host-side mirror of ScanShape (queries per wave).
inline int scan_qw(int dt, uint32_t ks) {
    const int lane = dt == FV_DTYPE_FP8_E4M3 ? 32 : 16;
    const int regs16 = (int)ks * (lane / 4);
    return 4 * regs16 <= 192 ? 64 : (2 * regs16 <= 192 ? 32 : 16);
}
inline int scan_row_bytes(int dt, uint32_t ks) {
    return (int)ks * 4 * (dt == FV_DTYPE_FP8_E4M3 ? 32 : 16);
}

template <typename V>
__device__ __forceinline__ V ld_vec(const uint8_t* p) {
    if constexpr (sizeof(V) == 16) {
        return __builtin_bit_cast(V, *reinterpret_cast<const uint4*>(p));
    } else {
        struct Two {
            uint4 a, b;
        } t;
        t.a = reinterpret_cast<const uint4*>(p)[0];
        t.b = reinterpret_cast<const uint4*>(p)[1];
        return __builtin_bit_cast(V, t);
    }
}

// Smallest float above f (finite f): a non-strict bound tested with the same
// strict compare as the chunk's own threshold.
__device__ __forceinline__ float next_up(float f) {
    uint32_t u = __float_as_uint(f);
    if (f == 0.0f) return __uint_as_float(1u);
    return __uint_as_float((u & 0x80000000u) ? u - 1u : u + 1u);
}

__device__ __forceinline__ float bound_from_ord(uint32_t g) {
    if (g == 0xFFFFFFFFu) return kAdmit;
    float f = unord_f32(g);
    return f < kAdmit ? next_up(f) : kAdmit;
}

// Shortlist depth of one query: factor * k (or the fixed shortlist), at least
// k, at most s_max.
__device__ __forceinline__ void query_k_s(const FlatSearchArgs& a, uint32_t q, uint32_t& k,
                                          uint32_t& S) {
    k = a.d_top_k ? a.d_top_k[q] : a.top_k;
    if (k > a.top_k) k = a.top_k;
    if (k == 0) k = 1;
    S = a.factor ? a.factor * k : (a.shortlist ? a.shortlist : 8u * k);
    if (S < k) S = k;
    if (S > a.s_max) S = a.s_max;
}

// BOUND = true is the first pass: same tiles and MMAs, but the epilogue only
// keeps a running minimum per (lane, element) — 16 disjoint row subsets per
// query per chunk — written to `minima`; rows are limited to the first
// `bound_rows` positions. No thresholds, buffers or prunes.
template <int DT, int METRIC, int KS, int R, bool BOUND, bool DMA>
__global__ __launch_bounds__(kMfmaThreads, (ScanShape<DT, KS>::kOcc)) void flat_mfma_scan_kernel(
    const uint8_t* __restrict__ X, uint32_t row_bytes, const float* __restrict__ norms,
    const float* __restrict__ row_scales, const uint8_t* __restrict__ Qh,
    const float* __restrict__ qscale, uint32_t nq, FlatSearchArgs a, uint32_t n_chunks,
    uint32_t qblocks, uint32_t cap, const uint32_t* __restrict__ rows,
    const uint32_t* __restrict__ nrows, uint32_t n, uint32_t* __restrict__ gthr,
    uint32_t* __restrict__ counts, uint64_t* __restrict__ cand, float* __restrict__ minima,
    uint32_t bound_rows) {
    using Ops = MfmaOps<DT>;
    using vec = typename Ops::vec;
    using Sh = ScanShape<DT, KS>;
    constexpr bool kFP8 = DT == FV_DTYPE_FP8_E4M3;
    constexpr int QW = Sh::kQW;
    constexpr int QB = Sh::kQB;
    constexpr int NJ = QW / 16;                  // 16-query column blocks per wave
    // Rows of 12 or more 16-byte fragment reads per lane stream their A
    // fragments (register budget: the query fragments take 192 VGPRs there).
    constexpr bool kStreamA = KS * (MfmaOps<DT>::kLaneBytes / 16) >= 12;
    // Rows per tile. The first pass (no dynamic LDS) and the workgroup-prune
    // instantiations keep their double-buffered tiles within 50 KB, so that
    // several workgroups share a CU, unless the registers allow only one
    // wave per SIMD anyway (then the taller tiles of the main pass).
    constexpr int RT = ((R == 0 || BOUND) && Sh::kOcc > 1) ? Sh::kRTDeep : Sh::kRT;
    constexpr int LB = Sh::kLaneBytes;
    constexpr int kRowBytes = Sh::kRowBytes;
    // Register staging: 16-bit rows padded by 32 bytes (conflict-free
    // fragment reads), FP8 rows by 16. Direct-to-LDS staging (DMA): rows
    // unpadded, 16-byte chunk c of row r stored at chunk c ^ (r & 15) of its
    // row (conflict-free for the same reads; rows are whole multiples of 16
    // chunks).
    constexpr int kLdsRow = DMA ? kRowBytes : kRowBytes + (kFP8 ? 16 : 32);
    constexpr int kThrPerRow = kMfmaThreads / RT;
    constexpr int kSegBytes = kRowBytes / kThrPerRow;  // bytes one thread stages per row
    constexpr int NPRE = DMA ? 1 : kSegBytes / 16;
    static_assert(kSegBytes % 16 == 0, "staging segments are whole 16-byte loads");
    static_assert(!DMA || kRowBytes % 256 == 0, "direct-to-LDS tiles need 256-byte rows");
    // Direct-to-LDS loads per wave per tile (1 KB each, tile bytes split over
    // the 4 waves).
    constexpr int kDmaPerWave = DMA ? RT * kRowBytes / 4096 : 1;
    static_assert(!DMA || (RT * kRowBytes) % 4096 == 0, "tiles split into 1 KB loads per wave");

    // Two tile buffers as distinct arrays, addressed with compile-time indices
    // (the tile loop is unrolled by two): the compiler can then tell that the
    // direct-to-LDS loads into one buffer do not alias the reads of the other
    // and waits for them only at the barrier.
    __shared__ __attribute__((aligned(16))) uint8_t s_tile0[RT * kLdsRow];
    __shared__ __attribute__((aligned(16))) uint8_t s_tile1[RT * kLdsRow];
    __shared__ __attribute__((aligned(16))) float s_cinit[2][RT];
    // Row scales: FP8 and cosine only (one slot otherwise, never read).
    constexpr bool kScaled = kFP8 || METRIC == METRIC_COSINE;
    __shared__ __attribute__((aligned(16))) float s_scale[2][kScaled ? RT : 1];
    // Row ids of the last four tiles (slot = tile index mod 4): records of a
    // tile are flushed up to one tile later, while other waves already store
    // the ids of the tile after next.
    __shared__ uint32_t s_rowid[4][RT];
    __shared__ uint32_t s_cnt[QB];
    __shared__ uint32_t s_keep[QB];
    __shared__ float s_thr[QB];
    // Queries to prune, one bit each; three sets used in turn by successive
    // checks (workgroup-prune scans only: the wave-prune scans fill the
    // 160 KB of LDS with two workgroups, to the byte at 128-dim FP16 rows).
    __shared__ __attribute__((aligned(16))) uint32_t s_need[3][R == 0 ? QB / 32 : 1];
    // Per-query candidate buffers (workgroup-prune scans): start and keys
    // of each query's buffer within the chunk.
    __shared__ uint32_t s_cbase[R == 0 ? QB : 1];
    __shared__ uint32_t s_ccap[R == 0 ? QB : 1];
    extern __shared__ uint64_t s_scratch[];
    auto tile_ptr = [&](auto bc) -> uint8_t* {
        if constexpr (decltype(bc)::value == 0) return s_tile0;
        else return s_tile1;
    };
    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int l16 = lane & 15;
    const int grp = lane >> 4;

    // Rows of the (possibly compacted) row space, and the chunks that cover
    // them: the launched count, or fewer when it is chosen from the device
    // row count (then the surplus workgroups zero their queries' counts).
    uint32_t n_eff = nrows ? *nrows : n;
    if (BOUND && n_eff > bound_rows) n_eff = bound_rows;
    const uint32_t b = blockIdx.x;
    if (!BOUND && a.chunk_want) {
        const uint32_t nch = chunks_for_rows(n_eff, a.chunk_want, a.chunk_min_rows, n_chunks);
        if (b >= nch * qblocks) {
            const uint32_t e = b - nch * qblocks;
            const uint32_t c = nch + e / qblocks, q = (e % qblocks) * QB + (uint32_t)threadIdx.x;
            if (threadIdx.x < QB && q < nq) counts[(size_t)c * nq + q] = 0u;
            return;
        }
        n_chunks = nch;
    }
    // XCD-first work mapping: consecutive block ids land on different XCDs;
    // within an XCD the chunk varies slowest.
    uint32_t chunk, qblock;
    if ((n_chunks & 7u) == 0u) {
        const uint32_t cpx = n_chunks >> 3;
        const uint32_t xcd = b & 7u, i = b >> 3;
        chunk = xcd * cpx + i / qblocks;
        qblock = i % qblocks;
    } else {
        chunk = b / qblocks;
        qblock = b % qblocks;
    }
    uint32_t rpc = (n_eff + n_chunks - 1) / n_chunks;
    rpc = (rpc + RT - 1) / RT * RT;
    const uint32_t r_begin = min(n_eff, chunk * rpc);
    const uint32_t r_end = min(n_eff, r_begin + rpc);
    const uint32_t qb0 = qblock * QB;
    const uint32_t qw0 = qb0 + wave * QW;
    constexpr bool kBlockPrune = R == 0;
    // Wave-prune instantiations: each wave's record staging.
    uint64_t* my_scratch = s_scratch + (size_t)wave * (kRecBytes / 8);
    BlockSelectLds& sel = *reinterpret_cast<BlockSelectLds*>(s_scratch);  // workgroup prunes

    // B fragments: lane (l16, grp) holds query qw0 + 16j + l16, bytes
    // LB*(4s + grp) .. +LB of the encoded row for step s (K chunks
    // interleaved over the four lane groups; the same permutation of K is
    // used for the A fragments, so every dim is paired exactly once).
    // Queries arrive pre-encoded (negated, rows of kRowBytes, padded to whole
    // query blocks).
    vec bfrag[NJ][KS];
    float qsc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const uint8_t* qrow = Qh + (size_t)(qw0 + 16 * j + l16) * kRowBytes + LB * grp;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            bfrag[j][s] = ld_vec<vec>(qrow + LB * 4 * s);
            // One wave per SIMD (large query fragments): keep the query
            // fragments in accumulation registers (AGPRs). The MMAs read their
            // B operand from there, and the architectural VGPRs stay free for
            // the row fragments, addresses and selection.
            if constexpr (Sh::kOcc == 1) __asm__ volatile("" : "+a"(bfrag[j][s]));
        }
        qsc[j] = kFP8 ? qscale[qw0 + 16 * j + l16] : 1.0f;
    }
    // Retire the fragment loads here: otherwise the wait-count tracking keeps
    // them "in flight" around the tile loop and every tile waits for its own
    // prefetches (vmcnt(0)) before its last MMAs.
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    // Per-query state; lane l < QW of wave w owns local query w*QW + l.
    const bool owner = lane < QW;
    const uint32_t my_ql = (uint32_t)(wave * QW + lane);
    const uint32_t my_q = qb0 + my_ql;
    const bool my_live = owner && my_q < nq;
    // Candidate buffers: per-query ones (cand_off, workgroup prunes only)
    // or `cap` keys per query.
    const bool percap = kBlockPrune && a.cand_off != nullptr;
    uint64_t* const cand_c = cand + (size_t)chunk * (percap ? a.cand_stride : (uint64_t)nq * cap);
    if (owner) {
        uint32_t k, S;
        if (my_live) query_k_s(a, my_q, k, S);
        else S = 1;
        s_keep[my_ql] = S;
        s_cnt[my_ql] = 0;
        float th = my_live
            ? bound_from_ord(__hip_atomic_load(&gthr[my_q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            : -__builtin_inff();
        if constexpr (kBlockPrune) {
            if (percap) {
                const uint32_t b0 = my_live ? a.cand_off[my_q] : 0u;
                const uint32_t nk = my_live ? a.cand_off[my_q + 1] - b0 : 0u;
                s_cbase[my_ql] = b0;
                s_ccap[my_ql] = nk;
                // A buffer cut short by the planned total (understated k
                // sum) keeps no candidates: the query returns no rows.
                const uint32_t smax_q = a.s_max < cap ? a.s_max : cap;
                if (nk < S + (cap - smax_q)) th = -__builtin_inff();
            }
        }
        s_thr[my_ql] = th;
    }
    if constexpr (kBlockPrune)
        if (tid < QB / 32) s_need[0][tid] = s_need[1][tid] = s_need[2][tid] = 0u;
    // Buffer of local query ql and its capacity in keys.
    auto qbuf = [&](uint32_t ql) -> uint64_t* {
        if constexpr (kBlockPrune)
            if (percap) return cand_c + s_cbase[ql];
        return cand_c + (size_t)(qb0 + ql) * cap;
    };
    auto qcap = [&](uint32_t ql) -> uint32_t {
        if constexpr (kBlockPrune)
            if (percap) return s_ccap[ql];
        return cap;
    };

    const uint32_t ntiles = (r_end - r_begin + RT - 1) / RT;
    // Tiles between buffer checks: after a prune a buffer holds at most
    // S + 64 keys (bisection slack), so pk more tiles fit in cap.
    const uint32_t smax_c = a.s_max < cap ? a.s_max : cap;
    uint32_t pk = cap > smax_c + 64u + RT ? (cap - smax_c - 64u) / RT : 1u;
    // Wide rows (few row tiles per tile, long tiles): with an even check
    // interval, records are flushed every second tile (the checks fall on
    // flushed tiles); otherwise after every tile.
    const bool flush_pairs = kStreamA && pk >= 2u;
    if (flush_pairs) pk &= ~1u;
    const int st_row = tid / kThrPerRow;
    const int st_seg = tid % kThrPerRow;
    // Staging registers of the next tile. Nothing loaded here is used until
    // store_tile, after this tile's MMAs, so the loads overlap the compute.
    uint4 pre[NPRE];
    uint32_t pre_row = 0, pre_mword = 0xFFFFFFFFu;
    float pre_norm = 0.0f, pre_rs = 1.0f;
    bool pre_in = false;
    // Compacted row space: row ids are fetched one tile ahead of the data.
    // All prefetch loads are unconditional from clamped, valid addresses (rows
    // past the chunk end get C = +inf, so their data never matters): no
    // exec-masked loads, so no conservative waits around them.
    uint32_t rid_next = 0;
    // Direct-to-LDS: load i of this lane covers LDS bytes
    // (wave * kDmaPerWave + i) * 1024 + 16 * lane of the tile, i.e. row
    // dma_row[i], global byte offset dma_off[i] within that row (the swizzle
    // applied in reverse); its row id is fetched a tile ahead like rid_next,
    // and its byte offset in the table when the tile's loads start.
    uint32_t dma_row[kDmaPerWave], dma_off[kDmaPerWave], dma_rid[kDmaPerWave];
    uint32_t dma_voff[kDmaPerWave];
    const uint32_t wave_u = __builtin_amdgcn_readfirstlane((uint32_t)wave);
    // The table as one buffer (the launch takes this path only for tables of
    // less than 4 GiB, so every row offset fits the 32-bit buffer offset).
    const __amdgpu_buffer_rsrc_t xrsrc = raw_rsrc(X, DMA ? n * row_bytes : 0u);
#pragma unroll
    for (int i = 0; i < kDmaPerWave; ++i) {
        const uint32_t o = (uint32_t)((wave * kDmaPerWave + i) * 1024 + 16 * lane);
        dma_row[i] = o / (uint32_t)kRowBytes;
        const uint32_t unit = (o % (uint32_t)kRowBytes) / 16u;
        dma_off[i] = (unit ^ (dma_row[i] & 15u)) * 16u;
        dma_rid[i] = 0;
        dma_voff[i] = 0;
    }
    auto fetch_rid = [&](uint32_t p0) {
        const uint32_t p = p0 + st_row;
        rid_next = rows ? rows[p < r_end ? p : r_end - 1] : p;
        if constexpr (DMA) {
#pragma unroll
            for (int i = 0; i < kDmaPerWave; ++i) {
                const uint32_t q = p0 + dma_row[i];
                dma_rid[i] = rows ? rows[q < r_end ? q : r_end - 1] : (q < r_end ? q : r_end - 1);
            }
        }
    };
    // Wide rows issue the next tile's direct-to-LDS loads one by one between
    // the MMAs of the first row tile (mma_stream); others at the tile start.
    constexpr bool kDmaSpread = DMA && kStreamA;
    constexpr int kDmaFirst = kDmaSpread ? 0 : kDmaPerWave;
    auto dma_rest = [&](auto dbc) {
#pragma unroll
        for (int i = kDmaFirst; i < kDmaPerWave; ++i)
            lds_dma16(xrsrc, dma_voff[i], tile_ptr(dbc) + (wave_u * kDmaPerWave + i) * 1024);
    };
    auto load_tile = [&](uint32_t p0, auto dbc) {
        const uint32_t p = p0 + st_row;
        pre_in = p < r_end;
        pre_row = pre_in ? rid_next : (rows ? rows[r_begin] : r_begin);
        if constexpr (DMA) {
            // Rows past the chunk end load a valid row (their C is +inf).
            // The four waves' loads queue on the CU's one memory pipeline,
            // and a wave stalls until its loads are taken, so wide rows spread
            // them over their MMAs. The offsets are computed here, once the
            // row ids have arrived (a later write of a register that a
            // direct-to-LDS load reads would wait for that load).
#pragma unroll
            for (int i = 0; i < kDmaPerWave; ++i) dma_voff[i] = dma_rid[i] * row_bytes + dma_off[i];
#pragma unroll
            for (int i = 0; i < kDmaFirst; ++i)
                lds_dma16(xrsrc, dma_voff[i], tile_ptr(dbc) + (wave_u * kDmaPerWave + i) * 1024);
        } else {
            const uint8_t* src = X + (size_t)pre_row * row_bytes;
#pragma unroll
            for (int i = 0; i < NPRE; ++i) {
                const uint32_t e = (uint32_t)(st_seg * kSegBytes + 16 * i);
                if (e + 16 <= row_bytes) pre[i] = *reinterpret_cast<const uint4*>(src + e);
                else pre[i] = make_uint4(0, 0, 0, 0);
            }
        }
        if (METRIC != METRIC_INNER_PRODUCT) pre_norm = norms[pre_row];
        if (kFP8) pre_rs = row_scales[pre_row];
        pre_mword = (a.mask && !rows) ? a.mask[pre_row >> 5] : 0xFFFFFFFFu;
    };
    auto store_tile = [&](auto bc, uint32_t slot) {
        constexpr int buf = decltype(bc)::value;
        if constexpr (!DMA) {
            uint8_t* dst = tile_ptr(bc) + st_row * kLdsRow + st_seg * kSegBytes;
#pragma unroll
            for (int i = 0; i < NPRE; ++i) *reinterpret_cast<uint4*>(dst + 16 * i) = pre[i];
        }
        if (st_seg == 0) {
            const bool ok = pre_in && ((pre_mword >> (pre_row & 31u)) & 1u);
            float c = __builtin_inff(), sc = 1.0f;
            if (ok) {
                if (METRIC == METRIC_L2) c = 0.5f * pre_norm;
                else c = 0.0f;
                if (METRIC == METRIC_COSINE) sc = pre_norm > 0.0f ? 1.0f / sqrtf(pre_norm) : 0.0f;
                if (kFP8) sc = sc * pre_rs;
            }
            s_cinit[buf][st_row] = c;
            if constexpr (kScaled) s_scale[buf][st_row] = sc;
            s_rowid[slot][st_row] = pre_row;
        }
    };

    // Wave-private staging of candidate records, flushed into the per-query
    // buffers by all 64 lanes at once: the scan loop only ballots the
    // per-lane minimum of 4 scores and writes one record per passing lane;
    // the per-score threshold tests and key building happen at the flush,
    // only for recorded lanes.
    // Wave mode: the records alias the wave's prune scratch (never live at the
    // same time). Workgroup mode: they follow the selection's LDS block.
    uint8_t* rec_base = kBlockPrune
        ? reinterpret_cast<uint8_t*>(s_scratch) + kScanSelBytes + (size_t)wave * kRecBytes
        : reinterpret_cast<uint8_t*>(my_scratch);
    floatx4* rec_score = reinterpret_cast<floatx4*>(rec_base);
    uint32_t* rec_meta = reinterpret_cast<uint32_t*>(rec_base + (size_t)kRec * 16);
    uint32_t nrec = 0;
    auto flush_records = [&]() {
        wave_lds_sync();
        for (uint32_t e = lane; e < nrec; e += kWave) {
            const uint32_t meta = rec_meta[e];
            const uint32_t ql = wave * QW + (meta & 0xFFu);
            const uint32_t rbase = ((meta >> 8) & 0xFFu);  // first of the 4 rows within the tile
            const uint32_t rslot = meta >> 16;             // row-id slot of the record's tile
            const float th = s_thr[ql];
            const floatx4 sc = rec_score[e];
            uint64_t* bq = qbuf(ql);
            // One LDS atomic per record (its passing scores together), then
            // the key stores: no store is followed by an LDS atomic of the
            // same record, so none has to complete before the next atomic.
            bool ok[4];
            uint32_t npass = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ok[i] = sc[i] < th;
                npass += ok[i] ? 1u : 0u;
            }
            if (npass) {
                uint32_t pos = atomicAdd(&s_cnt[ql], npass);
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (ok[i]) bq[pos++] = make_key(sc[i], s_rowid[rslot][rbase + i]);
            }
        }
        nrec = 0;
        wave_lds_sync();
    };

    if (ntiles > 0) {
        fetch_rid(r_begin);
        load_tile(r_begin, B0{});
        if constexpr (DMA) dma_rest(B0{});
        if (ntiles > 1) fetch_rid(r_begin + RT);
        store_tile(B0{}, 0u);
    }
    // Every wave's direct-to-LDS loads of the tile have landed before any
    // wave passes the barrier (the compiler's own waits cover only the
    // issuing wave's later LDS reads).
    if constexpr (DMA) __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();

    floatx4 runmin[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) runmin[j] = floatx4{kAdmit, kAdmit, kAdmit, kAdmit};

    // One tile: compute on buffer `buf` while the next tile is staged into
    // the other one.
    // Tile of the next buffer check, and the set of prune bits it uses.
    uint32_t next_check = pk - 1u;
    uint32_t ck = 1u;
    auto step = [&](uint32_t t, auto bc) {
        constexpr int buf = decltype(bc)::value;
        using Next = std::integral_constant<int, buf ^ 1>;
        const uint32_t p0 = r_begin + t * RT;
        // This tile's thresholds, C inputs and scales go to registers before
        // the next tile's loads are issued: an LDS read after a direct-to-LDS
        // load that the compiler cannot tell apart from its destination waits
        // for that load, which would serialize the prefetch.
        float thr[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) thr[j] = s_thr[wave * QW + 16 * j + l16];
        constexpr int NS = RT / 16;
        floatx4 cin_t[NS], scl_t[NS];
#pragma unroll
        for (int rt = 0; rt < NS; ++rt) {
            cin_t[rt] = *reinterpret_cast<const floatx4*>(&s_cinit[buf][rt * 16 + 4 * grp]);
            if (kFP8 || METRIC == METRIC_COSINE)
                scl_t[rt] = *reinterpret_cast<const floatx4*>(&s_scale[buf][rt * 16 + 4 * grp]);
        }
        if (t + 1 < ntiles) {
            load_tile(p0 + RT, Next{});
            if (!kStreamA && t + 2 < ntiles) fetch_rid(p0 + 2 * RT);
        }
        // Prefetch the shared bound; it is consumed after this tile's MMAs.
        const uint32_t gpre =
            __hip_atomic_load(&gthr[my_live ? my_q : 0u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

        // Row tile rt (16 rows): MMAs into acc; scores after the FP8 / cosine
        // scaling; then selection. The MMAs of tile rt+1 are issued before
        // the selection of tile rt, so that VALU work overlaps the matrix
        // pipe.
        // A fragments of row tile rt (16 rows) from LDS.
        auto load_frags = [&](int rt, vec* afrag) {
            if constexpr (DMA) {
                // Row rt*16 + l16 (so its swizzle is l16); the lane's K bytes
                // LB*(4s + grp) are chunks (LB/16)*(4s + grp) (+1).
                const uint8_t* rb = tile_ptr(bc) + (rt * 16 + l16) * kRowBytes;
                constexpr int CPL = LB / 16;
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const uint32_t u = (uint32_t)(CPL * (4 * s + grp));
                    if constexpr (CPL == 1) {
                        afrag[s] = ld_vec<vec>(rb + ((u ^ (uint32_t)l16) * 16u));
                    } else {
                        struct Two {
                            uint4 a, b;
                        } t2;
                        t2.a = *reinterpret_cast<const uint4*>(rb + ((u ^ (uint32_t)l16) * 16u));
                        t2.b = *reinterpret_cast<const uint4*>(rb + (((u + 1u) ^ (uint32_t)l16) * 16u));
                        afrag[s] = __builtin_bit_cast(vec, t2);
                    }
                }
            } else {
                const uint8_t* arow = tile_ptr(bc) + (rt * 16 + l16) * kLdsRow + grp * LB;
#pragma unroll
                for (int s = 0; s < KS; ++s) afrag[s] = ld_vec<vec>(arow + LB * 4 * s);
            }
        };
        auto load_frag_step = [&](int rt, int s) -> vec {
            if constexpr (DMA) {
                const uint8_t* rb = tile_ptr(bc) + (rt * 16 + l16) * kRowBytes;
                constexpr int CPL = LB / 16;
                const uint32_t u = (uint32_t)(CPL * (4 * s + grp));
                if constexpr (CPL == 1) {
                    return ld_vec<vec>(rb + ((u ^ (uint32_t)l16) * 16u));
                } else {
                    struct Two {
                        uint4 a, b;
                    } t2;
                    t2.a = *reinterpret_cast<const uint4*>(rb + ((u ^ (uint32_t)l16) * 16u));
                    t2.b = *reinterpret_cast<const uint4*>(rb + (((u + 1u) ^ (uint32_t)l16) * 16u));
                    return __builtin_bit_cast(vec, t2);
                }
            } else {
                const uint8_t* arow = tile_ptr(bc) + (rt * 16 + l16) * kLdsRow + grp * LB;
                return ld_vec<vec>(arow + LB * 4 * s);
            }
        };
        auto mma_frags = [&](int rt, const vec* afrag, floatx4* acc) {
            const floatx4 cin = cin_t[rt];
#pragma unroll
            for (int j = 0; j < NJ; ++j) acc[j] = kFP8 ? floatx4{0.0f, 0.0f, 0.0f, 0.0f} : cin;
            // NJ independent accumulation chains, interleaved.
#pragma unroll
            for (int s = 0; s < KS; ++s)
#pragma unroll
                for (int j = 0; j < NJ; ++j) acc[j] = Ops::mma(afrag[s], bfrag[j][s], acc[j]);
        };
        // A fragments of row tile rt read kAhead steps ahead of their MMAs
        // through a register ring that runs on into the next row tile (its
        // first steps are read during this tile's last MMAs); the scheduling
        // groups keep that order (one LDS read, then the step's NJ MMAs),
        // which the scheduler would otherwise collapse into read-wait-MMA
        // pairs.
        constexpr int kAhead = KS < 4 ? KS : 4;
        constexpr int kReads = MfmaOps<DT>::kLaneBytes / 16;  // ds_read_b128 per fragment
        // Loads of the next tile, or zero-size ones into the idle buffer
        // after the last tile (keeps the MMA stream free of branches).
        const __amdgpu_buffer_rsrc_t dma_rs = t + 1 < ntiles ? xrsrc : raw_rsrc(X, 0u);
        auto mma_stream = [&](int rt, floatx4* acc, vec* ring) {
            constexpr int D = kDmaSpread ? kDmaPerWave : 0;
            const floatx4 cin = cin_t[rt];
#pragma unroll
            for (int j = 0; j < NJ; ++j) acc[j] = kFP8 ? floatx4{0.0f, 0.0f, 0.0f, 0.0f} : cin;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                // Ring slots follow the running step index rt*KS + s.
                const int slot = (rt * KS + s) % kAhead;
                const vec a = ring[slot];
                const bool more = s + kAhead < KS || rt + 1 < NS;
                if (s + kAhead < KS) ring[slot] = load_frag_step(rt, s + kAhead);
                else if (rt + 1 < NS) ring[slot] = load_frag_step(rt + 1, s + kAhead - KS);
#pragma unroll
                for (int j = 0; j < NJ; ++j) acc[j] = Ops::mma(a, bfrag[j][s], acc[j]);
                if (more) __builtin_amdgcn_sched_group_barrier(0x100, kReads, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, NJ, 0);
                if (rt == 0 && D > 0) {
                    // Loads i in [s*D/KS, (s+1)*D/KS), fenced off from the
                    // scheduler so they stay between the steps' MMAs.
                    const int i0 = s * D / KS, i1 = (s + 1) * D / KS;
#pragma unroll
                    for (int i = 0; i < D; ++i) {
                        if (i >= i0 && i < i1) {
                            __builtin_amdgcn_sched_barrier(0);
                            lds_dma16(dma_rs, dma_voff[i],
                                      tile_ptr(Next{}) + (wave_u * kDmaPerWave + i) * 1024);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    }
                }
            }
        };
        auto select_tile = [&](int rt, floatx4* acc) {
            if (kFP8 || METRIC == METRIC_COSINE) {
                const floatx4 scl = scl_t[rt];
                const floatx4 cin = cin_t[rt];
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    if (kFP8) acc[j] = cin + acc[j] * (scl * qsc[j]);
                    else acc[j] = acc[j] * scl;
                }
            }
            if constexpr (BOUND) {
#pragma unroll
                for (int j = 0; j < NJ; ++j)
#pragma unroll
                    for (int i = 0; i < 4; ++i) runmin[j][i] = fminf(runmin[j][i], acc[j][i]);
                return;
            }
            // One compare per lane and query column on the minimum of its 4
            // scores; passing lanes write a record (compacted wave-wide with
            // ballot + mbcnt). Columns of queries past nq have threshold -inf.
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const float m4 = fminf(fminf(acc[j][0], acc[j][1]), fminf(acc[j][2], acc[j][3]));
                const bool pass = m4 < thr[j];
                const uint64_t bm = __ballot(pass);
                if (bm == 0) continue;
                if (pass) {
                    const uint32_t idx = nrec + __builtin_amdgcn_mbcnt_hi(
                        (uint32_t)(bm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u));
                    rec_score[idx] = acc[j];
                    rec_meta[idx] = (uint32_t)(16 * j + l16) | ((uint32_t)(rt * 16 + 4 * grp) << 8) |
                                    ((t & 3u) << 16);
                }
                nrec += (uint32_t)__popcll(bm);
                if (nrec > kRec - kWave) flush_records();
            }
        };
        floatx4 acc0[NJ], acc1[NJ];
        if constexpr (kStreamA) {
            // Wide rows: the A fragments of a row tile are read step by step
            // next to their MMAs (no whole-tile fragment sets in registers);
            // the MMAs of rt+1 are still issued before the selection of rt.
            vec ring[kAhead];
#pragma unroll
            for (int s = 0; s < kAhead; ++s) ring[s] = load_frag_step(0, s);
            __builtin_amdgcn_sched_group_barrier(0x100, kAhead * kReads, 0);
            mma_stream(0, acc0, ring);
            if (t + 2 < ntiles) fetch_rid(p0 + 2 * RT);
#pragma unroll
            for (int rt = 0; rt < NS; ++rt) {
                floatx4* cur = (rt & 1) ? acc1 : acc0;
                floatx4* nxt = (rt & 1) ? acc0 : acc1;
                if (rt + 1 < NS) mma_stream(rt + 1, nxt, ring);
                select_tile(rt, cur);
            }
        } else {
            // Software pipeline over the row tiles of this tile: the fragments
            // of row tile rt+2 are read from LDS while the MMAs of rt+1 run and
            // the selection of rt is done, so no MMA waits on its LDS read.
            vec fr0[KS], fr1[KS];
            load_frags(0, fr0);
            if (NS > 1) load_frags(1, fr1);
            mma_frags(0, fr0, acc0);
#pragma unroll
            for (int rt = 0; rt < NS; ++rt) {
                floatx4* cur = (rt & 1) ? acc1 : acc0;
                floatx4* nxt = (rt & 1) ? acc0 : acc1;
                const vec* fn = (rt & 1) ? fr0 : fr1;  // row tile rt+1, loaded a step ahead
                vec* fnn = (rt & 1) ? fr1 : fr0;       // free: row tile rt's MMAs are issued
                if (rt + 1 < NS) mma_frags(rt + 1, fn, nxt);
                if (rt + 2 < NS) load_frags(rt + 2, fnn);
                select_tile(rt, cur);
            }
        }
        // A pair of tiles shares one flush (the row ids of both stay in
        // their slots until two tiles later).
        if (!BOUND && (buf == 1 || !flush_pairs)) flush_records();
        // Workgroup prunes: at a check, each owner marks its query before the
        // tile barrier (its count is final once its wave has flushed), so the
        // check needs no barrier of its own. Check i uses set i % 3 and
        // clears set (i + 1) % 3, last read by check i - 2, before the
        // barrier that preceded check i - 1.
        const bool check = t == next_check;
        if constexpr (kBlockPrune && !BOUND) {
            if (check) {
                if (tid < QB / 32) s_need[ck == 2u ? 0u : ck + 1u][tid] = 0u;
                if (my_live && s_cnt[my_ql] + pk * RT > qcap(my_ql))
                    atomicOr(&s_need[ck][my_ql >> 5], 1u << (my_ql & 31));
            }
        }
        if (t + 1 < ntiles) store_tile(Next{}, (t + 1) & 3u);
        if constexpr (DMA) __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        __syncthreads();
        if (BOUND) return;

        // A buffer gains at most one key per row, so buffers are checked
        // every pk tiles and pruned when pk more tiles could overflow them.
        if (!check) {
            if (my_live) s_thr[my_ql] = fminf(s_thr[my_ql], bound_from_ord(gpre));
            wave_lds_sync();
            return;
        }
        next_check += pk;
        if (!kBlockPrune) {
            // Wave-private prunes of buffers that could overflow on the next tile.
            const uint32_t c = my_live ? s_cnt[my_ql] : 0u;
            uint64_t need = __ballot(my_live && c + pk * RT > cap);
            while (need) {
                const int bl = __builtin_ctzll(need);
                need &= need - 1;
                const uint32_t qlb = wave * QW + bl;
                const uint32_t qgb = qb0 + qlb;
                const uint32_t keep = s_keep[qlb];
                uint64_t* bq = qbuf(qlb);
                uint32_t th;
                uint32_t kept = wave_bisect_prune<(R > 0 ? R : 1)>(bq, s_cnt[qlb], keep, lane, &th);
                if (kept + pk * RT > cap) {
                    // Heavy ties at the bound: exact (distance, row) selection.
                    wave_lds_sync();
                    kept = wave_key_prune<(R > 0 ? R : 1)>(bq, kept, keep, lane, &th);
                }
                if (lane == 0) {
                    s_cnt[qlb] = kept;
                    if (th != 0xFFFFFFFFu) {
                        s_thr[qlb] = fminf(s_thr[qlb], unord_f32(th));
                        atomicMin(&gthr[qgb], th);
                    }
                }
                wave_lds_sync();
            }
        } else {
            // Workgroup prunes, one query at a time: radix selection of the
            // `keep` smallest keys in place in the query's buffer (global
            // memory, L2-resident), so no buffer-sized LDS scratch is needed.
            // The queries were marked in set ck before this tile's barrier.
            // Most checks find none: the words are read together (one LDS
            // round trip, not one per word) and the loop is skipped.
            uint32_t any = 0;
#pragma unroll
            for (int w = 0; w < QB / 32; ++w) any |= s_need[ck][w];
            const int w_end = __builtin_amdgcn_readfirstlane(any) != 0u ? QB / 32 : 0;
            for (int w = 0; w < w_end; ++w) {
                uint32_t bits = s_need[ck][w];
                while (bits) {
                    const int bl = __builtin_ctz(bits);
                    bits &= bits - 1;
                    const uint32_t ql = w * 32 + bl;
                    const uint32_t qg = qb0 + ql;
                    const uint32_t c = s_cnt[ql];
                    const uint32_t keep = s_keep[ql];
                    uint64_t* bq = qbuf(ql);
                    // The trigger implies c > cap_q - pk * RT >= S_q + 64 > keep;
                    // the keys of one buffer are distinct (one per row).
                    if (c > keep) {
                        const uint64_t kth = block_select(bq, c, keep, sel);
                        if (tid == 0) {
                            s_cnt[ql] = keep;
                            const float th = key_dist(kth);
                            s_thr[ql] = fminf(s_thr[ql], th);
                            atomicMin(&gthr[qg], ord_f32(th));
                        }
                    }
                    __syncthreads();
                }
            }
            ck = ck == 2u ? 0u : ck + 1u;
        }
        // Tighten thresholds with the prefetched shared bound. Each wave only
        // reads its own queries' thresholds, so a wave-level sync suffices.
        if (my_live) s_thr[my_ql] = fminf(s_thr[my_ql], bound_from_ord(gpre));
        wave_lds_sync();
    };
    for (uint32_t t = 0; t < ntiles; t += 2) {
        step(t, B0{});
        if (t + 1 < ntiles) step(t + 1, B1{});
    }
    if (BOUND) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const uint32_t qg = qw0 + 16 * j + l16;
            if (qg < nq) {
                float* mo = minima + ((size_t)chunk * nq + qg) * 16 + 4 * grp;
#pragma unroll
                for (int i = 0; i < 4; ++i) mo[i] = runmin[j][i];
            }
        }
        return;
    }
    flush_records();  // an odd last tile
    if (my_live) counts[(size_t)chunk * nq + my_q] = s_cnt[my_ql];
}

// Static LDS of one main-pass instantiation (queried once); with the plan's
// dynamic LDS it must fit the 160 KB of a CU.
template <int DT, int METRIC, int KS, int R, bool DMA>
static size_t scan_static_lds() {
    static const size_t v = [] {
        hipFuncAttributes attr{};
        const void* f = reinterpret_cast<const void*>(flat_mfma_scan_kernel<DT, METRIC, KS, R, false, DMA>);
        return hipFuncGetAttributes(&attr, f) == hipSuccess ? attr.sharedSizeBytes : (size_t)1 << 30;
    }();
    return v;
}

template <int DT, int METRIC, int KS, int R, bool DMA>
static hipError_t launch_scan_r(const ScanLaunch& L) {
    const FlatMfmaPlan& p = *L.p;
    const dim3 g(p.qblocks * p.n_chunks);
    flat_mfma_scan_kernel<DT, METRIC, KS, R, false, DMA><<<g, kMfmaThreads, p.lds_scan, L.stream>>>(
        L.X, L.row_bytes, L.norms, L.row_scales, L.Qh, L.qscale, L.nq, L.a, p.n_chunks, p.qblocks,
        p.cap, L.rows, L.nrows, L.n, L.gthr, L.counts, L.cand, nullptr, 0u);
    return hipSuccess;
}

template <int DT, int METRIC, int KS, bool DMA>
static hipError_t launch_scan_d(const ScanLaunch& L) {
    const FlatMfmaPlan& p = *L.p;
    const size_t st = p.prune_regs == 16 ? scan_static_lds<DT, METRIC, KS, 16, DMA>()
                                         : scan_static_lds<DT, METRIC, KS, 0, DMA>();
    if (st + p.lds_scan > 160u * 1024u) return hipErrorInvalidValue;
    if (p.bound_chunks) {
        const dim3 g(p.qblocks * p.bound_chunks);
        flat_mfma_scan_kernel<DT, METRIC, KS, 16, true, DMA><<<g, kMfmaThreads, 0, L.stream>>>(
            L.X, L.row_bytes, L.norms, L.row_scales, L.Qh, L.qscale, L.nq, L.a, p.bound_chunks,
            p.qblocks, p.cap, L.rows, L.nrows, L.n, L.gthr, L.counts, L.cand, L.minima,
            p.bound_rows);
        launch_bound_select(L);
    }
    // The kernel timer brackets the main pass only.
    void* timer = timer_begin(L.stream);
    const hipError_t e = p.prune_regs == 16 ? launch_scan_r<DT, METRIC, KS, 16, DMA>(L)
                                            : launch_scan_r<DT, METRIC, KS, 0, DMA>(L);
    timer_end(timer, L.stream);
    return e;
}

// Direct-to-LDS staging when the rows are whole 256-byte multiples with no
// K padding beyond the table's own (row stride == staged row bytes) and the
// table spans less than 4 GiB (32-bit buffer offsets).
template <int DT, int METRIC, int KS>
static hipError_t launch_scan(const ScanLaunch& L) {
    if constexpr (scan_dma_ok<DT, KS>()) {
        if (L.row_bytes == (uint32_t)ScanShape<DT, KS>::kRowBytes &&
            (uint64_t)L.n * L.row_bytes <= 0xFFFFFFFFull)
            return launch_scan_d<DT, METRIC, KS, true>(L);
    }
    return launch_scan_d<DT, METRIC, KS, false>(L);
}

template <int DT, int KS>
static hipError_t launch_scan_metric(const ScanLaunch& L) {
    if (L.a.metric == METRIC_L2) return launch_scan<DT, METRIC_L2, KS>(L);
    if (L.a.metric == METRIC_INNER_PRODUCT) return launch_scan<DT, METRIC_INNER_PRODUCT, KS>(L);
    return launch_scan<DT, METRIC_COSINE, KS>(L);
}

}  // namespace fv
