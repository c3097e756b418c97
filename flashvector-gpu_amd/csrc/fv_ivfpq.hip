// IVF-PQ ADC list scan fused with shortlist selection and exact FP32 rerank
// (cuda_ivf_pq_search_batch and fv_ivf_pq_search).
//
// Reference: kernels/src/ivf_pq_lookup.cu:14-475 with the coarse quantizer of
// kernels/src/ivf_rerank.cuh:49-130. The reference keeps per-thread candidate
// lists merged serially (shortlist <= 32) or appends under a racy cutoff with
// global-memory bitonic prunes, and reads each entry's codes by vector id.
//
// Here, per query:
//  * ivf_coarse_kernel: every centroid distance as an fmaf chain over d (the
//    reference's order), then the nprobe smallest by (distance, cluster) from
//    one LDS bitonic sort.
//  * ivf_scan_kernel (one workgroup per query; the hardware dispatcher hands
//    queries of uneven cost to free CUs as they drain):
//      for each probed list: the residual LUT [M][256] is built in LDS (fmaf
//      chain over the sub-dimension, the reference's order), then the list is
//      streamed 256 entries per pass. With a list-ordered code copy every
//      lane reads its entry's 16 code bytes with one coalesced 128-bit load;
//      with the reference's id-indexed codes it gathers them by vector id.
//      ADC = sequential FP32 adds over m = 0..M-1, abandoned after any
//      fourth term once the sum is past the threshold's distance (LUT
//      entries are >= 0, so such an entry could not be kept; kept entries
//      add every term). An entry is kept if its
//      (adc, id) key is below the running threshold key; keys go to a
//      candidate buffer (LDS, or global memory for very deep shortlists)
//      that is cut to the S best by radix selection whenever the next keys
//      would not fit.
//    finally the S best are reranked with the exact FP32 distance (the
//    reference CPU path's order), cut to the k best and written in
//    (distance, id) order.
// No MFMA: the scan is bound by its LUT lookups in LDS and by list bytes.
#include "fv_ivfpq.hpp"
#include "fv_workspace.hpp"

namespace fv {

constexpr int kIvfThreads = 256;
constexpr uint32_t kPassesPerCheck = 4;  // 256-entry passes between capacity checks
constexpr uint32_t kEntriesPerCheck = kIvfThreads * kPassesPerCheck;

__global__ __launch_bounds__(256) void ivf_coarse_kernel(
    const float* __restrict__ C, uint32_t nlist, uint32_t dim, const float* __restrict__ Qm,
    const uint32_t* __restrict__ nprobe_q, uint32_t nprobe_u, uint32_t max_nprobe, uint32_t npow,
    uint32_t* __restrict__ probes) {
    extern __shared__ uint64_t s_ck[];
    const uint32_t q = blockIdx.x;
    const float* qv = Qm + (size_t)q * dim;
    for (uint32_t c = threadIdx.x; c < npow; c += blockDim.x) {
        uint64_t key = kEmptyKey;
        if (c < nlist) {
            const float* cen = C + (size_t)c * dim;
            float dist = 0.0f;
            for (uint32_t d = 0; d < dim; ++d) {
                float diff = qv[d] - cen[d];
                dist = __builtin_fmaf(diff, diff, dist);
            }
            if (dist < kAdmit) key = make_key(dist, c);
        }
        s_ck[c] = key;
    }
    __syncthreads();
    block_bitonic_sort(s_ck, npow);
    uint32_t np = nprobe_q ? nprobe_q[q] : nprobe_u;
    if (np == 0) np = 1;
    if (np > nlist) np = nlist;
    for (uint32_t i = threadIdx.x; i < max_nprobe; i += blockDim.x) {
        uint64_t k = i < np ? s_ck[i] : kEmptyKey;
        probes[(size_t)q * max_nprobe + i] = k == kEmptyKey ? FV_EMPTY_ID : key_slot(k);
    }
}

// Coarse probe for lists too many to sort in LDS: every centroid's key
// (distance as in ivf_coarse_kernel, cluster id; keys of unadmitted
// distances have all-ones distance bits and stay distinct) goes to the
// query's row of `keys`, and a radix selection keeps the np smallest there.
// The probes come out in no particular order: the scan's result depends only
// on the set of probed lists.
__global__ __launch_bounds__(kSelThreads) void ivf_coarse_global_kernel(
    const float* __restrict__ C, uint32_t nlist, uint32_t dim, const float* __restrict__ Qm,
    const uint32_t* __restrict__ nprobe_q, uint32_t nprobe_u, uint32_t max_nprobe,
    uint64_t* __restrict__ keys, uint32_t* __restrict__ probes) {
    __shared__ BlockSelectLds L;
    const uint32_t q = blockIdx.x;
    const float* qv = Qm + (size_t)q * dim;
    uint64_t* buf = keys + (size_t)q * nlist;
    for (uint32_t c = threadIdx.x; c < nlist; c += blockDim.x) {
        const float* cen = C + (size_t)c * dim;
        float dist = 0.0f;
        for (uint32_t d = 0; d < dim; ++d) {
            float diff = qv[d] - cen[d];
            dist = __builtin_fmaf(diff, diff, dist);
        }
        buf[c] = dist < kAdmit ? make_key(dist, c) : (0xFFFFFFFF00000000ull | c);
    }
    __syncthreads();
    uint32_t np = nprobe_q ? nprobe_q[q] : nprobe_u;
    if (np == 0) np = 1;
    if (np > nlist) np = nlist;
    if (np > max_nprobe) np = max_nprobe;  // the probe row holds max_nprobe slots
    if (np < nlist) (void)block_select(buf, nlist, np, L);
    // Admitted probes packed at the front (the scan stops at the first empty
    // one), empty slots after them.
    __shared__ uint32_t s_n;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    uint32_t* pq = probes + (size_t)q * max_nprobe;
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
        const uint64_t k = buf[i];
        if ((k >> 32) != 0xFFFFFFFFull) pq[atomicAdd(&s_n, 1u)] = key_slot(k);
    }
    __syncthreads();
    for (uint32_t i = s_n + threadIdx.x; i < max_nprobe; i += blockDim.x) pq[i] = FV_EMPTY_ID;
}

// Workgroup control words and selection scratch of the scan kernel, in
// dynamic LDS after the LUT and the candidate buffer.
struct IvfCtl {
    BlockSelectLds bs;                         // block_select scratch
    uint32_t item;                             // position of the current query in `order`
    uint32_t wn[3][kIvfThreads / kWave];       // per-wave keys to append: 2 checks, after a cut
};
constexpr size_t kIvfCtlBytes = (sizeof(IvfCtl) + 15) / 16 * 16;
// Distance bits of an admitted key are never all ones (ord of a NaN).
constexpr uint64_t kUnadmittedKey = 0xFFFFFFFF00000000ull;
static_assert(kIvfThreads == (int)kSelThreads, "block_select and one thread per PQ codeword");

// LUT [M][256] of one probed list: lut[m][j] = sum over s of
// (res[m*SUB+s] - codebook[m][j][s])^2 as an fmaf chain in s. One codeword
// per thread; the codebook rows of NM sub-quantizers (8 float4 loads per
// thread) are issued together before their sums.
template <int UPM>
__device__ __forceinline__ void build_lut_vec(float* lut, const float* __restrict__ cbk,
                                              const float* res, uint32_t M) {
    constexpr uint32_t SUB = 4 * UPM, NM = 8 / UPM;
    const uint32_t j = threadIdx.x;
    for (uint32_t m0 = 0; m0 < M; m0 += NM) {
        float4 v[NM][UPM];
#pragma unroll
        for (uint32_t a = 0; a < NM; ++a)
#pragma unroll
            for (uint32_t u = 0; u < (uint32_t)UPM; ++u)
                if (m0 + a < M)
                    v[a][u] = reinterpret_cast<const float4*>(cbk + ((size_t)(m0 + a) * 256u + j) * SUB)[u];
#pragma unroll
        for (uint32_t a = 0; a < NM; ++a) {
            if (m0 + a < M) {
                const float* r = res + (m0 + a) * SUB;
                float dist = 0.0f, diff;
#pragma unroll
                for (uint32_t u = 0; u < (uint32_t)UPM; ++u) {
                    diff = r[4 * u + 0] - v[a][u].x; dist = __builtin_fmaf(diff, diff, dist);
                    diff = r[4 * u + 1] - v[a][u].y; dist = __builtin_fmaf(diff, diff, dist);
                    diff = r[4 * u + 2] - v[a][u].z; dist = __builtin_fmaf(diff, diff, dist);
                    diff = r[4 * u + 3] - v[a][u].w; dist = __builtin_fmaf(diff, diff, dist);
                }
                lut[(m0 + a) * 256u + j] = dist;
            }
        }
    }
}

__device__ __forceinline__ void build_lut(float* lut, const float* __restrict__ cbk, const float* res,
                                          uint32_t M, uint32_t sub) {
    const bool vec = (reinterpret_cast<uintptr_t>(cbk) & 15u) == 0u;
    if (vec && sub == 8u) {
        build_lut_vec<2>(lut, cbk, res, M);
    } else if (vec && sub == 4u) {
        build_lut_vec<1>(lut, cbk, res, M);
    } else if (vec && sub == 16u) {
        build_lut_vec<4>(lut, cbk, res, M);
    } else if (vec && sub == 32u) {
        build_lut_vec<8>(lut, cbk, res, M);
    } else {
        const uint32_t j = threadIdx.x;
        for (uint32_t m = 0; m < M; ++m) {
            const float* cb = cbk + ((size_t)m * 256u + j) * sub;
            const float* r = res + m * sub;
            float dist = 0.0f;
            for (uint32_t s = 0; s < sub; ++s) {
                float diff = r[s] - cb[s];
                dist = __builtin_fmaf(diff, diff, dist);
            }
            lut[m * 256u + j] = dist;
        }
    }
}

// Arguments of the scan kernel.
struct IvfScanArgs {
    IvfPqGpuTables t;
    const uint8_t* codes_lo;      // list-ordered codes, or nullptr (gather by vector id)
    const float* Qm;
    const uint32_t* probes;       // [nq][max_nprobe], packed, FV_EMPTY_ID after the last
    uint32_t max_nprobe;
    const uint32_t* k_q;
    uint32_t k_u;
    uint32_t shortlist;           // S of the plan (>= every query's S); partial slot size
    uint32_t cap;
    int do_rerank;
    const uint32_t* mask;
    uint64_t* gbuf;
    GpuSearchResult* out;
    uint32_t out_stride;
    // Work items: query order[i] owns items item_start[i] .. item_start[i+1]-1,
    // each a contiguous run of its probed entries (lists in probe order).
    uint32_t nq;
    const uint32_t* order;        // queries, heaviest first
    const uint32_t* item_start;   // [nq + 1]
    const uint32_t* cost;         // [nq] probed entries per query
    const uint32_t* slot_start;   // [nq] first partial slot of a split query
    uint32_t* next_item;
    uint32_t* done;               // [nq] finished items of a split query
    uint64_t* qbound;             // [nq] shared admission bound of a split query
    uint64_t* part_keys;          // [slots][shortlist] item shortlists of split queries
    uint32_t* part_cnt;           // [slots]
};

// Persistent: each workgroup repeatedly takes the next work item through a
// global atomic counter. A query whose probed lists hold more entries than
// the batch's chunk length is split into items of at most that many entries
// (its lists concatenated in probe order), so one query over a huge (Zipf
// head) list is scanned by several workgroups at once; queries are ordered
// heaviest first. Each item of a split query keeps its S best keys in a
// partial slot and publishes its S-th best to the query's shared bound (any
// item's S-th best bounds the query's S-th best, so the other items admit
// only keys at or below it); the item that finishes last merges the slots
// and finishes the query. The result depends only on the set of probed
// entries, so it is the same however the query is split.
//
// M16 (16 sub-quantizers): the LUT is a static LDS array, the kernel's only
// static LDS object and so at LDS address 0, which lets each lookup's byte
// extract and scale fold into one SDWA shift (the address is code*4 plus an
// immediate). Other M keep the LUT at the start of dynamic LDS. With
// list-ordered codes the ids and codes of the next check are loaded (buffer
// loads, zero past the run's end) while the current one is scored.
//
// Candidates: each check of kEntriesPerCheck entries computes its keys into
// registers and every wave publishes how many are below the threshold. Only
// if appending them would overflow the buffer is it cut to the S best
// (block_select) and the threshold lowered (and the waves recount). Keys are
// then appended at positions from the per-wave counts, so the buffer fill and
// the threshold are uniform registers, identical in every thread.
template <bool LISTORDER, bool GBUF, int METRIC, bool M16>
__global__ __launch_bounds__(kIvfThreads, 4) void ivf_scan_kernel(IvfScanArgs a) {
    // Dynamic LDS: LUT [M][256] f32 (unless static), candidate buffer (unless
    // in global memory), control words, the probed list's residual [dim].
    __shared__ float s_lut16[M16 ? 16 * 256 : 1];
    __shared__ float s_lmin[M16 ? 16 : 1];
    extern __shared__ uint8_t s_dyn[];
    const IvfPqGpuTables& t = a.t;
    const uint32_t cap = a.cap;
    float* lut = M16 ? s_lut16 : reinterpret_cast<float*>(s_dyn);
    const size_t lut_bytes = M16 ? 0 : (size_t)t.m_pq * 256 * sizeof(float);
    uint64_t* buf = GBUF ? a.gbuf + (size_t)blockIdx.x * cap : reinterpret_cast<uint64_t*>(s_dyn + lut_bytes);
    uint8_t* ctl_base = s_dyn + lut_bytes + (GBUF ? (size_t)0 : (size_t)cap * 8);
    IvfCtl& ctl = *reinterpret_cast<IvfCtl*>(ctl_base);
    float* res = reinterpret_cast<float*>(ctl_base + kIvfCtlBytes);
    constexpr uint32_t kW = kIvfThreads / kWave;
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const uint32_t n_items = a.item_start[a.nq];

  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) ctl.item = atomicAdd(a.next_item, 1u);
    __syncthreads();
    const uint32_t item = __builtin_amdgcn_readfirstlane(ctl.item);
    if (item >= n_items) return;
    // The item's query: the last position i with item_start[i] <= item.
    uint32_t lo = 0, hi = a.nq;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.item_start[mid] <= item) lo = mid;
        else hi = mid;
    }
    const uint32_t q = a.order[lo];
    const uint32_t nit = a.item_start[lo + 1] - a.item_start[lo], part = item - a.item_start[lo];
    const uint32_t T = a.cost[q];
    const uint64_t clen = ((uint64_t)T + nit - 1u) / nit;
    const uint32_t e0 = (uint32_t)min((uint64_t)T, part * clen);
    const uint32_t e1 = (uint32_t)min((uint64_t)T, e0 + clen);
    const bool split = nit > 1u;

    const uint32_t M = t.m_pq, sub = t.sub_dim, dim = t.dim;
    const float* qv = a.Qm + (size_t)q * dim;
    uint32_t k = a.k_q ? a.k_q[q] : a.k_u;
    if (k > a.out_stride) k = a.out_stride;
    uint32_t S = a.shortlist < k ? k : a.shortlist;
    if (S > cap - kEntriesPerCheck) S = cap - kEntriesPerCheck;
    if (S == 0) S = 1;
    uint32_t cnt = 0;          // keys in the candidate buffer
    uint64_t thr = kEmptyKey;  // admission threshold: the S-th best key so far
    uint32_t ph = 0;           // which wn[] row the current check uses
    bool publish = split;      // publish cuts to the query's shared bound

    // Appends the keys of one check that are below `lim` (<= thr), cutting
    // the buffer to the S best first if they would not fit.
    auto append = [&](const uint64_t (&kk)[kPassesPerCheck], uint64_t lim) {
        // wn[ph] row: written before this check's barrier, read after it
        // and before the next one; the next write to it is two checks on.
        uint32_t wn = 0;
#pragma unroll
        for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass)
            wn += (uint32_t)__popcll(__ballot(kk[pass] < lim));
        if (lane == 0) ctl.wn[ph][wave] = wn;
        __syncthreads();
        uint32_t row = ph, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < kW; ++w) total += ctl.wn[ph][w];
        if (cnt + total > cap) {
            thr = uniform_u64(block_select(buf, cnt, S, ctl.bs));
            cnt = S;
            if (publish && threadIdx.x == 0) atomicMin((unsigned long long*)&a.qbound[q], thr);
            lim = lim < thr ? lim : thr;
            wn = 0;
#pragma unroll
            for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass)
                wn += (uint32_t)__popcll(__ballot(kk[pass] < lim));
            if (lane == 0) ctl.wn[2][wave] = wn;
            __syncthreads();
            row = 2;
            total = 0;
#pragma unroll
            for (uint32_t w = 0; w < kW; ++w) total += ctl.wn[2][w];
        }
        if (total != 0u) {
            uint32_t pos = cnt;
#pragma unroll
            for (uint32_t w = 0; w < kW; ++w) pos += w < wave ? ctl.wn[row][w] : 0u;
#pragma unroll
            for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass) {
                const bool in = kk[pass] < lim;
                const uint64_t bal = __ballot(in);
                const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                           __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                if (in) buf[pos + below] = kk[pass];
                pos += (uint32_t)__popcll(bal);
            }
        }
        cnt += total;
        ph ^= 1u;
    };
    // Admission bound of the next check: the own threshold, and for a split
    // query also its shared bound (keys are distinct, so key <= bound is
    // key < bound + 1). The shared bound is read one check ahead, always (an
    // unconditional load keeps the wait counts of the prefetched loads exact).
    uint64_t gnext = kEmptyKey;
    auto bound = [&]() -> uint64_t {
        const uint64_t g = uniform_u64(gnext);
        gnext = __hip_atomic_load(&a.qbound[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!publish) return thr;
        const uint64_t gb = g == kEmptyKey ? kEmptyKey : g + 1u;
        return gb < thr ? gb : thr;
    };

    // Keys of 16-sub-quantizer entries: four lookups per segment; an entry
    // whose running sum plus a lower bound of its remaining terms is past the
    // bound's distance skips the rest; admitted entries add all 16 terms in
    // order. LUT entries are sums of squares (>= 0 or NaN), so the rounded
    // running sum never decreases and, rounding being monotone, the final sum
    // is at least the sequential sum of the partial sum and the remaining
    // sub-quantizers' LUT minima, itself at least (partial + exact sum of
    // those minima) * (1 - 12u). lmin[k] under-estimates that exact suffix
    // sum and thr_hi = thr * (1 + 2^-18) covers the relative errors (at most
    // 13u with u = 2^-24), so an entry is only abandoned when its final sum
    // exceeds the bound: results equal those of the plain sequential sums.
    float lmin[3] = {0.0f, 0.0f, 0.0f};
    auto keys_m16 = [&](uint32_t n, const uint32_t (&ids)[kPassesPerCheck],
                        const uint4 (&cw)[kPassesPerCheck], float thr_d, bool masked,
                        uint64_t (&kk)[kPassesPerCheck]) {
        const float thr_hi = thr_d * (1.0f + 0x1p-18f);
        auto seg4 = [&](int seg, uint32_t w, float s) {
            float l[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) l[b] = s_lut16[(seg * 4 + b) * 256u + ((w >> (8 * b)) & 255u)];
            // The first segment starts from its first term: 0 + x == x for
            // every LUT entry (never -0: fmaf chains of squares from +0).
            s = seg == 0 ? l[0] : s + l[0];
#pragma unroll
            for (int b = 1; b < 4; ++b) s = s + l[b];
            // Issue the four lookups together (addresses, LDS reads, adds),
            // not one round trip at a time.
            __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);
            return s;
        };
#pragma unroll
        for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass) {
            kk[pass] = kEmptyKey;
            if (pass * kIvfThreads + threadIdx.x < n && (!masked || mask_allows(a.mask, ids[pass]))) {
                float s = seg4(0, cw[pass].x, 0.0f);
                if (s + lmin[0] <= thr_hi) {
                    s = seg4(1, cw[pass].y, s);
                    if (s + lmin[1] <= thr_hi) {
                        s = seg4(2, cw[pass].z, s);
                        if (s + lmin[2] <= thr_hi) {
                            s = seg4(3, cw[pass].w, s);
                            if (s <= thr_d && s < kAdmit) kk[pass] = make_key(s, ids[pass]);
                        }
                    }
                }
            }
        }
    };
    auto dist_of = [](uint64_t lim) {
        return lim == kEmptyKey ? __builtin_inff() : key_dist(lim);
    };
    // List-ordered ids and codes of the check at list position `base` (none
    // at or past `end`: the loads then return zeros without reading memory).
    auto load_lo = [&](uint32_t base, uint32_t end, uint32_t (&ids)[kPassesPerCheck],
                       uint4 (&cw)[kPassesPerCheck]) {
        // base and end are uniform; readfirstlane keeps the descriptors in
        // scalar registers.
        base = __builtin_amdgcn_readfirstlane(base);
        const uint32_t left = base < end ? end - base : 0u;
        const uint32_t n = __builtin_amdgcn_readfirstlane(left < kEntriesPerCheck ? left : kEntriesPerCheck);
        const __amdgpu_buffer_rsrc_t rc = raw_rsrc(a.codes_lo + (size_t)base * 16u, n * 16u);
        const __amdgpu_buffer_rsrc_t ri = raw_rsrc(t.d_ivf_vec_ids + base, n * 4u);
#pragma unroll
        for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass) {
            const uint32_t e = pass * kIvfThreads + threadIdx.x;
            cw[pass] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rc, e * 16u, 0, 0));
            ids[pass] = __builtin_amdgcn_raw_buffer_load_b32(ri, e * 4u, 0, 0);
        }
    };
    // Scores list positions [beg, end) of the list whose LUT is built.
    auto scan_run = [&](uint32_t beg, uint32_t end) {
        if constexpr (M16 && LISTORDER) {
            if (a.mask == nullptr) {
                // Two register sets, alternating: the next check's loads are
                // issued (unconditionally) before the current one is scored.
                uint32_t idA[kPassesPerCheck], idB[kPassesPerCheck];
                uint4 cwA[kPassesPerCheck], cwB[kPassesPerCheck];
                uint64_t kk[kPassesPerCheck];
                load_lo(beg, end, idA, cwA);
                for (uint32_t base = beg; base < end; base += 2u * kEntriesPerCheck) {
                    const uint32_t b1 = base + kEntriesPerCheck;
                    load_lo(b1, end, idB, cwB);
                    uint64_t lim = bound();
                    keys_m16(end - base, idA, cwA, dist_of(lim), false, kk);
                    append(kk, lim);
                    if (b1 >= end) break;
                    load_lo(b1 + kEntriesPerCheck, end, idA, cwA);
                    lim = bound();
                    keys_m16(end - b1, idB, cwB, dist_of(lim), false, kk);
                    append(kk, lim);
                }
                return;
            }
        }
        for (uint32_t base = beg; base < end; base += kEntriesPerCheck) {
            const uint64_t lim = bound();
            const float thr_d = dist_of(lim);
            uint64_t kk[kPassesPerCheck];
            if constexpr (M16) {
                uint32_t ids[kPassesPerCheck];
                uint4 cw[kPassesPerCheck];
                if constexpr (LISTORDER) {
                    load_lo(base, end, ids, cw);
                } else {
                    // Codes gathered by vector id.
#pragma unroll
                    for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass) {
                        const uint32_t e = base + pass * kIvfThreads + threadIdx.x;
                        ids[pass] = t.d_ivf_vec_ids[e < end ? e : end - 1];
                    }
#pragma unroll
                    for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass) {
                        const uint8_t* cp = t.d_pq_codes + (size_t)ids[pass] * 16u;
                        if ((reinterpret_cast<uintptr_t>(cp) & 15u) == 0u) {
                            cw[pass] = *reinterpret_cast<const uint4*>(cp);
                        } else {
                            uint32_t w4[4];
                            for (int b = 0; b < 4; ++b)
                                w4[b] = cp[4 * b] | (cp[4 * b + 1] << 8) | (cp[4 * b + 2] << 16) | ((uint32_t)cp[4 * b + 3] << 24);
                            cw[pass] = make_uint4(w4[0], w4[1], w4[2], w4[3]);
                        }
                    }
                }
                keys_m16(end - base, ids, cw, thr_d, true, kk);
            } else {
#pragma unroll
                for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass) {
                    const uint32_t e = base + pass * kIvfThreads + threadIdx.x;
                    kk[pass] = kEmptyKey;
                    if (e >= end) continue;
                    const uint32_t id = t.d_ivf_vec_ids[e];
                    if (!mask_allows(a.mask, id)) continue;
                    const uint8_t* cp = LISTORDER ? a.codes_lo + (size_t)e * M : t.d_pq_codes + (size_t)id * M;
                    float adc = 0.0f;
                    bool full = true;
                    for (uint32_t m = 0; m < M; ++m) {
                        adc = adc + lut[m * 256u + cp[m]];
                        if ((m & 3u) == 3u && m + 1u < M && !(adc <= thr_d)) {
                            full = false;
                            break;
                        }
                    }
                    if (full && adc < kAdmit) kk[pass] = make_key(adc, id);
                }
            }
            append(kk, lim);
        }
    };

    // This item's entries: positions [e0, e1) of the concatenation of the
    // probed lists.
    uint64_t cum = 0;
    for (uint32_t p = 0; p < a.max_nprobe && cum < e1; ++p) {
        const uint32_t c = a.probes[(size_t)q * a.max_nprobe + p];
        if (c >= t.nlist) break;  // probes are packed; the rest are empty
        const uint32_t lb = t.d_ivf_offsets[c], le = t.d_ivf_offsets[c + 1];
        const uint64_t len = le - lb;
        const uint64_t s0 = cum > e0 ? cum : e0, s1 = cum + len < e1 ? cum + len : e1;
        if (s0 < s1) {
            const float* cen = t.d_centroids + (size_t)c * dim;
            // The previous list's LUT build is done with res (barrier after it).
            for (uint32_t d = threadIdx.x; d < dim; d += kIvfThreads) res[d] = qv[d] - cen[d];
            __syncthreads();  // res visible; the previous list's lookups are done with the LUT
            build_lut(lut, t.d_pq_codebooks, res, M, sub);
            __syncthreads();
            if constexpr (M16) {
                static_assert(kIvfThreads == 256, "16 threads per sub-quantizer table");
                // Minimum of each sub-quantizer's 256 entries (16 threads per
                // table, conflict-free strided reads, then a 16-lane
                // reduction), and the lower bounds of the last 12 / 8 / 4
                // terms.
                const uint32_t m = threadIdx.x >> 4, r = threadIdx.x & 15u;
                float v = lut[m * 256u + r];
#pragma unroll
                for (uint32_t i = 1; i < 16; ++i) v = fminf(v, lut[m * 256u + i * 16u + r]);
#pragma unroll
                for (int o = 8; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 16));
                if (r == 0) s_lmin[m] = v;
                __syncthreads();
                float suf = 0.0f;
#pragma unroll
                for (int k = 15; k >= 0; --k) {
                    suf = suf + s_lmin[k];
                    if ((k & 3) == 0 && k > 0) lmin[k / 4 - 1] = suf * (1.0f - 0x1p-19f);
                }
                // No entry of the list can reach the own threshold (which only
                // falls): skip its scan. The threshold is the same in every
                // wave (the shared bound is not: each wave reads it itself),
                // so the whole workgroup skips together.
                const float all = suf * (1.0f - 0x1p-19f);
                if (all > dist_of(thr) * (1.0f + 0x1p-18f)) {
                    cum += len;
                    continue;
                }
            }
            scan_run(lb + (uint32_t)(s0 - cum), lb + (uint32_t)(s1 - cum));
        }
        cum += len;
    }
    __syncthreads();
    uint32_t kept = cnt;
    if (kept > S) {
        block_select(buf, kept, S, ctl.bs);
        kept = S;
        __syncthreads();
    }
    if (split) {
        // Keep this item's shortlist; the last item of the query merges.
        const uint32_t slot = a.slot_start[q] + part;
        uint64_t* dst = a.part_keys + (size_t)slot * a.shortlist;
        for (uint32_t i = threadIdx.x; i < kept; i += kIvfThreads) dst[i] = buf[i];
        if (threadIdx.x == 0) a.part_cnt[slot] = kept;
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) ctl.item = atomicAdd(&a.done[q], 1u) + 1u == nit ? 1u : 0u;
        __syncthreads();
        if (__builtin_amdgcn_readfirstlane(ctl.item) == 0u) continue;
        __threadfence();
        cnt = 0;
        thr = kEmptyKey;
        publish = false;
        const uint32_t s0 = a.slot_start[q];
        for (uint32_t sl = s0; sl < s0 + nit; ++sl) {
            const uint32_t n = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&a.part_cnt[sl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            uint64_t* src = a.part_keys + (size_t)sl * a.shortlist;
            for (uint32_t base = 0; base < n; base += kEntriesPerCheck) {
                uint64_t kk[kPassesPerCheck];
#pragma unroll
                for (uint32_t pass = 0; pass < kPassesPerCheck; ++pass) {
                    const uint32_t e = base + pass * kIvfThreads + threadIdx.x;
                    kk[pass] = e < n ? __hip_atomic_load(&src[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                     : kEmptyKey;
                }
                append(kk, thr);
            }
        }
        __syncthreads();
        kept = cnt;
        if (kept > S) {
            block_select(buf, kept, S, ctl.bs);
            kept = S;
            __syncthreads();
        }
    }
    // Exact rerank of the shortlist, in place. Rows that are not admitted get
    // distinct keys above every admitted one (distance bits all ones).
    if (a.do_rerank && t.d_vectors) {
        for (uint32_t i = threadIdx.x; i < kept; i += blockDim.x) {
            const uint32_t id = key_slot(buf[i]);
            const float d = id < t.num_vectors
                ? exact_distance_vec<METRIC>(qv, t.d_vectors + (size_t)id * dim, dim)
                : FV_EMPTY_DIST;
            buf[i] = d < kAdmit ? make_key(d, id) : (kUnadmittedKey | i);
        }
        __syncthreads();
    }
    // The k best, in (distance, id) order: cut to k, then each key is written
    // at its rank among the k (or, for k above one key per thread, sorted).
    const uint32_t kout = k < kept ? k : kept;
    if (kept > kout && kout > 0u) {
        block_select(buf, kept, kout, ctl.bs);
        __syncthreads();
    }
    if (kout <= (uint32_t)kIvfThreads) {
        const uint64_t mine = threadIdx.x < kout ? buf[threadIdx.x] : kEmptyKey;
        uint32_t rank = 0;
        if (threadIdx.x < kout)
            for (uint32_t j = 0; j < kout; ++j) rank += buf[j] < mine ? 1u : 0u;
        __syncthreads();
        if (threadIdx.x < kout) buf[rank] = mine;
        __syncthreads();
    } else {
        const uint32_t n = next_pow2(kout);
        for (uint32_t i = kout + threadIdx.x; i < n; i += blockDim.x) buf[i] = kEmptyKey;
        __syncthreads();
        block_bitonic_sort(buf, n);
    }
    for (uint32_t i = threadIdx.x; i < a.out_stride; i += blockDim.x) {
        GpuSearchResult r;
        const uint64_t key = i < kout ? buf[i] : kEmptyKey;
        if ((key >> 32) == 0xFFFFFFFFull) {
            r.id = FV_EMPTY_ID;
            r.distance = FV_EMPTY_DIST;
        } else {
            r.id = key_slot(key);
            r.distance = key_dist(key);
        }
        a.out[(size_t)q * a.out_stride + i] = r;
    }
  }
}

// Scan cost of each query (entries over its probed lists), bucketed by
// log2 so that the heaviest queries are dispatched first. Costs saturate at
// 2^32 - 1 entries.
__global__ void ivf_cost_kernel(const uint32_t* __restrict__ probes, uint32_t max_nprobe,
                                const uint32_t* __restrict__ offsets, uint32_t nlist,
                                uint32_t nq, uint32_t* __restrict__ cost_out,
                                uint32_t* __restrict__ bucket, uint32_t* __restrict__ bucket_count) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint64_t cost = 0;
    for (uint32_t p = 0; p < max_nprobe; ++p) {
        const uint32_t c = probes[(size_t)q * max_nprobe + p];
        if (c >= nlist) break;
        cost += offsets[c + 1] - offsets[c];
    }
    const uint32_t c32 = cost > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)cost;
    cost_out[q] = c32;
    const uint32_t b = (uint32_t)__clz(c32 | 1u);  // 0 = heaviest
    bucket[q] = b;
    atomicAdd(&bucket_count[b], 1u);
}

__global__ void ivf_order_kernel(const uint32_t* __restrict__ bucket,
                                 const uint32_t* __restrict__ bucket_count, uint32_t nq,
                                 uint32_t* __restrict__ bucket_fill, uint32_t* __restrict__ order) {
    __shared__ uint32_t s_off[33];
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int b = 0; b < 33; ++b) {
            s_off[b] = run;
            run += bucket_count[b];
        }
    }
    __syncthreads();
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
        const uint32_t b = bucket[q];
        order[s_off[b] + atomicAdd(&bucket_fill[b], 1u)] = q;
    }
}

// Exclusive prefix sums of two per-thread counts over a 1024-thread
// workgroup; returns the totals in tot0/tot1.
__device__ __forceinline__ void block_scan2(uint32_t& v0, uint32_t& v1, uint32_t& tot0,
                                            uint32_t& tot1, uint32_t* s0, uint32_t* s1) {
    const uint32_t tid = threadIdx.x;
    s0[tid] = v0;
    s1[tid] = v1;
    __syncthreads();
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
        const uint32_t a0 = tid >= off ? s0[tid - off] : 0u, a1 = tid >= off ? s1[tid - off] : 0u;
        __syncthreads();
        s0[tid] += a0;
        s1[tid] += a1;
        __syncthreads();
    }
    tot0 = s0[blockDim.x - 1];
    tot1 = s1[blockDim.x - 1];
    v0 = s0[tid] - v0;
    v1 = s1[tid] - v1;
    __syncthreads();
}

// Work items of the batch (one workgroup of 1024 threads). Chunk length C =
// max(chunk_min, ceil(total / target_items), ceil(2 * total / max_slots)); a
// query of more than C entries becomes ceil(cost / C) items, every other
// query one item. The last bound keeps the split queries' items, at most
// 2 * total / C, within the max_slots partial slots. target_items = 0 or
// max_slots = 0 splits nothing.
__global__ __launch_bounds__(1024) void ivf_items_kernel(
    const uint32_t* __restrict__ cost, const uint32_t* __restrict__ order, uint32_t nq,
    uint32_t chunk_min, uint32_t target_items, uint32_t max_slots,
    uint32_t* __restrict__ item_start, uint32_t* __restrict__ slot_start) {
    __shared__ uint32_t s0[1024], s1[1024];
    __shared__ unsigned long long s_total;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_total = 0;
    __syncthreads();
    uint64_t my = 0;
    for (uint32_t i = tid; i < nq; i += blockDim.x) my += cost[i];
    atomicAdd(&s_total, (unsigned long long)my);
    __syncthreads();
    const uint64_t total = s_total;
    uint64_t C = ~0ull;
    if (target_items != 0u && max_slots != 0u) {
        C = chunk_min ? chunk_min : 1u;
        const uint64_t c1 = (total + target_items - 1u) / target_items;
        const uint64_t c2 = (2u * total + max_slots - 1u) / max_slots;
        C = C > c1 ? C : c1;
        C = C > c2 ? C : c2;
    }
    auto items_of = [&](uint32_t i) -> uint32_t {
        const uint64_t c = cost[order[i]];
        return c > C ? (uint32_t)((c + C - 1u) / C) : 1u;
    };
    // Each thread owns a contiguous run of positions.
    const uint32_t per = (nq + blockDim.x - 1u) / blockDim.x;
    const uint32_t i0 = min(nq, tid * per), i1 = min(nq, i0 + per);
    uint32_t ni = 0, ns = 0;
    for (uint32_t i = i0; i < i1; ++i) {
        const uint32_t n = items_of(i);
        ni += n;
        ns += n > 1u ? n : 0u;
    }
    uint32_t tot_i, tot_s;
    block_scan2(ni, ns, tot_i, tot_s, s0, s1);
    for (uint32_t i = i0; i < i1; ++i) {
        const uint32_t n = items_of(i);
        item_start[i] = ni;
        slot_start[order[i]] = ns;
        ni += n;
        ns += n > 1u ? n : 0u;
    }
    if (tid == 0) item_start[nq] = tot_i;
}

// Probe list of each query from the coarse top-k results: the first np_q
// cluster ids, empty past np_q.
__global__ void probes_from_results_kernel(const GpuSearchResult* __restrict__ res, uint32_t nq,
                                           uint32_t max_nprobe, const uint32_t* __restrict__ nprobe_q,
                                           uint32_t nprobe_u, uint32_t nlist,
                                           uint32_t* __restrict__ probes) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nq * max_nprobe) return;
    const uint32_t q = (uint32_t)(i / max_nprobe), p = (uint32_t)(i % max_nprobe);
    uint32_t np = nprobe_q ? nprobe_q[q] : nprobe_u;
    if (np == 0) np = 1;
    if (np > nlist) np = nlist;
    probes[i] = p < np ? res[i].id : FV_EMPTY_ID;
}

__global__ void list_order_codes_kernel(const uint8_t* __restrict__ codes,
                                        const uint32_t* __restrict__ vec_ids, uint32_t n,
                                        uint32_t M, uint8_t* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * M) return;
    size_t e = i / M, m = i % M;
    out[i] = codes[(size_t)vec_ids[e] * M + m];
}

IvfPqPlan plan_ivfpq(const IvfPqGpuTables* t, uint32_t nq, uint32_t max_nprobe,
                     uint32_t shortlist, uint32_t max_k) {
    IvfPqPlan p{};
    uint32_t S = shortlist < max_k ? max_k : shortlist;
    p.shortlist = S;
    p.max_nprobe = max_nprobe;
    // Room for S kept keys plus one check's keys, and for the final sort of
    // up to next_pow2(k) keys (k <= 1024).
    p.cap = (S + kEntriesPerCheck + 63u) / 64u * 64u;
    const uint32_t sort_room = next_pow2(S < 1024u ? S : 1024u);
    if (p.cap < sort_room) p.cap = sort_room;
    p.npow = next_pow2(t->nlist < 2u ? 2u : t->nlist);
    const size_t lut_bytes = (size_t)t->m_pq * 256 * sizeof(float);
    p.gbuf = lut_bytes + (size_t)p.cap * 8 > 64 * 1024;
    // Dynamic part only: with 16 sub-quantizers the LUT is static LDS.
    p.lds_scan = (t->m_pq == 16u ? 0 : lut_bytes) + (p.gbuf ? 0 : (size_t)p.cap * 8) +
                 kIvfCtlBytes + (size_t)t->dim * sizeof(float);
    p.grid = kMaxScanGrid;
    // Partial shortlists of split queries: at most kMaxPartBytes of slots.
    const size_t slot_bytes = (size_t)S * 8u;
    const size_t slots = kMaxPartBytes / slot_bytes;
    p.max_slots = (uint32_t)(slots < kMaxSlots ? slots : kMaxSlots);
    p.target_items = kDefaultTargetItems;
    p.chunk_min = kDefaultChunkMin;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    p.off_probes = take((size_t)nq * max_nprobe * 4);
    p.off_bucket = take((size_t)nq * 4);
    p.off_order = take((size_t)nq * 4);
    p.off_counters = take(68 * 4);  // 33 counts, 33 fills, next item, pad
    p.off_gbuf = take(p.gbuf ? (size_t)p.grid * p.cap * 8 : 0);
    p.off_cost = take((size_t)nq * 4);
    p.off_items = take(((size_t)nq + 1) * 4);
    p.off_slots = take((size_t)nq * 4);
    p.off_done = take((size_t)nq * 4);
    p.off_qbound = take((size_t)nq * 8);
    p.off_part_cnt = take((size_t)p.max_slots * 4);
    p.off_part_keys = take((size_t)p.max_slots * slot_bytes);
    if (max_nprobe <= kTiledCoarseMaxProbes && t->nlist >= 64u) p.coarse_mode = kCoarseTiled;
    else if (p.npow <= kLdsCoarseMaxLists) p.coarse_mode = kCoarseLds;
    else p.coarse_mode = kCoarseGlobal;
    if (p.coarse_mode == kCoarseTiled) {
        p.coarse_plan = plan_flat_exact(t->nlist, t->dim, nq, max_nprobe, METRIC_L2);
        p.off_coarse_res = take((size_t)nq * max_nprobe * sizeof(GpuSearchResult));
        p.off_coarse_ws = take(p.coarse_plan.workspace_bytes);
    } else if (p.coarse_mode == kCoarseGlobal) {
        p.off_coarse_keys = take((size_t)nq * t->nlist * sizeof(uint64_t));
    }
    p.workspace_bytes = off;
    return p;
}

template <bool LO, bool GB, bool M16>
static void launch_scan_m(int metric, dim3 g, size_t lds, hipStream_t s, const IvfScanArgs& a) {
    if (metric == METRIC_INNER_PRODUCT)
        ivf_scan_kernel<LO, GB, METRIC_INNER_PRODUCT, M16><<<g, kIvfThreads, lds, s>>>(a);
    else if (metric == METRIC_COSINE)
        ivf_scan_kernel<LO, GB, METRIC_COSINE, M16><<<g, kIvfThreads, lds, s>>>(a);
    else
        ivf_scan_kernel<LO, GB, METRIC_L2, M16><<<g, kIvfThreads, lds, s>>>(a);
}

template <bool LO, bool GB>
static void launch_scan_lo_gb(bool m16, int metric, dim3 g, size_t lds, hipStream_t s,
                              const IvfScanArgs& a) {
    if (m16) launch_scan_m<LO, GB, true>(metric, g, lds, s, a);
    else launch_scan_m<LO, GB, false>(metric, g, lds, s, a);
}

hipError_t launch_ivfpq(const IvfPqGpuTables* t, const uint8_t* codes_lo, const float* Qm,
                        uint32_t nq, const uint32_t* k_q, uint32_t k_u, const uint32_t* nprobe_q,
                        uint32_t nprobe_u, int do_rerank, int metric, const uint32_t* mask,
                        GpuSearchResult* out, uint32_t out_stride, void* workspace,
                        const IvfPqPlan& p, hipStream_t stream) {
    // LUT (static with 16 sub-quantizers) + dynamic LDS must fit one CU's LDS.
    if ((t->m_pq == 16u ? 16384u : 0u) + p.lds_scan > 160u * 1024u) return hipErrorInvalidValue;
    char* ws = static_cast<char*>(workspace);
    uint32_t* probes = reinterpret_cast<uint32_t*>(ws + p.off_probes);
    uint32_t* bucket = reinterpret_cast<uint32_t*>(ws + p.off_bucket);
    uint32_t* order = reinterpret_cast<uint32_t*>(ws + p.off_order);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(ws + p.off_counters);
    uint32_t* cost = reinterpret_cast<uint32_t*>(ws + p.off_cost);
    if (p.coarse_mode == kCoarseTiled) {
        // Tiled exact scan with fmaf accumulation: same values and (distance,
        // cluster) order as the per-query coarse kernel, at GEMM-like reuse.
        GpuSearchResult* cres = reinterpret_cast<GpuSearchResult*>(ws + p.off_coarse_res);
        const hipError_t ce = launch_flat_exact(t->d_centroids, nullptr, t->nlist, t->dim, Qm, nq,
                                                p.max_nprobe, METRIC_L2, nullptr, cres, p.max_nprobe,
                                                ws + p.off_coarse_ws, p.coarse_plan, stream, true);
        if (ce != hipSuccess) return ce;
        const size_t tot = (size_t)nq * p.max_nprobe;
        probes_from_results_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, stream>>>(
            cres, nq, p.max_nprobe, nprobe_q, nprobe_u, t->nlist, probes);
    } else if (p.coarse_mode == kCoarseLds) {
        ivf_coarse_kernel<<<nq, 256, (size_t)p.npow * 8, stream>>>(
            t->d_centroids, t->nlist, t->dim, Qm, nprobe_q, nprobe_u, p.max_nprobe, p.npow, probes);
    } else {
        ivf_coarse_global_kernel<<<nq, kSelThreads, 0, stream>>>(
            t->d_centroids, t->nlist, t->dim, Qm, nprobe_q, nprobe_u, p.max_nprobe,
            reinterpret_cast<uint64_t*>(ws + p.off_coarse_keys), probes);
    }
    // Counters, the split queries' finished-item counts and shared bounds
    // start cleared (fill kernels, so the sequence can be captured).
    fill_u32(ctr, 68, 0u, stream);
    fill_u32(reinterpret_cast<uint32_t*>(ws + p.off_done), nq, 0u, stream);
    fill_u32(reinterpret_cast<uint32_t*>(ws + p.off_qbound), (size_t)nq * 2, 0xFFFFFFFFu, stream);
    ivf_cost_kernel<<<(nq + 255) / 256, 256, 0, stream>>>(probes, p.max_nprobe, t->d_ivf_offsets,
                                                         t->nlist, nq, cost, bucket, ctr);
    ivf_order_kernel<<<(nq + 255) / 256 < 64 ? (nq + 255) / 256 : 64, 256, 0, stream>>>(
        bucket, ctr, nq, ctr + 33, order);
    uint32_t* item_start = reinterpret_cast<uint32_t*>(ws + p.off_items);
    uint32_t* slot_start = reinterpret_cast<uint32_t*>(ws + p.off_slots);
    ivf_items_kernel<<<1, 1024, 0, stream>>>(cost, order, nq, p.chunk_min, p.target_items,
                                             p.max_slots, item_start, slot_start);
    IvfScanArgs a;
    a.t = *t;
    a.codes_lo = codes_lo;
    a.Qm = Qm;
    a.probes = probes;
    a.max_nprobe = p.max_nprobe;
    a.k_q = k_q;
    a.k_u = k_u;
    a.shortlist = p.shortlist;
    a.cap = p.cap;
    a.do_rerank = do_rerank;
    a.mask = mask;
    a.gbuf = p.gbuf ? reinterpret_cast<uint64_t*>(ws + p.off_gbuf) : nullptr;
    a.out = out;
    a.out_stride = out_stride;
    a.nq = nq;
    a.order = order;
    a.item_start = item_start;
    a.cost = cost;
    a.slot_start = slot_start;
    a.next_item = ctr + 66;
    a.done = reinterpret_cast<uint32_t*>(ws + p.off_done);
    a.qbound = reinterpret_cast<uint64_t*>(ws + p.off_qbound);
    a.part_keys = reinterpret_cast<uint64_t*>(ws + p.off_part_keys);
    a.part_cnt = reinterpret_cast<uint32_t*>(ws + p.off_part_cnt);
    void* timer = timer_begin(stream);
    const dim3 g(p.grid);
    const bool lo = codes_lo != nullptr;
    const bool m16 = t->m_pq == 16u;
    if (lo && p.gbuf) launch_scan_lo_gb<true, true>(m16, metric, g, p.lds_scan, stream, a);
    else if (lo) launch_scan_lo_gb<true, false>(m16, metric, g, p.lds_scan, stream, a);
    else if (p.gbuf) launch_scan_lo_gb<false, true>(m16, metric, g, p.lds_scan, stream, a);
    else launch_scan_lo_gb<false, false>(m16, metric, g, p.lds_scan, stream, a);
    timer_end(timer, stream);
    return hipGetLastError();
}

static bool tables_ok(const IvfPqGpuTables* t) {
    return t->d_centroids && t->d_pq_codebooks && t->d_pq_codes && t->d_ivf_offsets &&
           t->d_ivf_vec_ids && t->dim > 0 && t->m_pq > 0 && t->sub_dim > 0 &&
           t->m_pq * t->sub_dim == t->dim;
}

}  // namespace fv

extern "C" {

// cuda_bridge.h:96-107 / kernels/src/ivf_pq_lookup.cu:339-473.
int cuda_ivf_pq_search_batch(const IvfPqGpuTables* tables, const float* d_queries,
                             uint32_t num_queries, uint32_t top_k, uint32_t nprobe,
                             uint32_t rerank_k, MetricType metric, const uint32_t* d_valid_mask,
                             GpuSearchResult* d_out_results, void* stream) {
    if (!tables || !d_queries || !d_out_results) return -1;
    if (num_queries == 0 || top_k == 0 || tables->m_pq == 0 || tables->nlist == 0) return 0;
    if (!fv::tables_ok(tables)) return (int)hipErrorInvalidValue;
    if (top_k > 256u) top_k = 256u;
    uint32_t np = nprobe == 0 ? 1u : nprobe;
    if (np > tables->nlist) np = tables->nlist;
    uint32_t S = rerank_k == 0 ? top_k : rerank_k;
    if (S < top_k) S = top_k;
    if (S > 10240u) S = 10240u;
    const int do_rerank = rerank_k != 1u && tables->d_vectors != nullptr;
    hipStream_t s = static_cast<hipStream_t>(stream);
    fv::IvfPqPlan p = fv::plan_ivfpq(tables, num_queries, np, S, top_k);
    fv::WorkspaceLease ws;
    hipError_t e = fv::stream_workspace(s, p.workspace_bytes, ws);
    if (e != hipSuccess) return (int)e;
    e = fv::launch_ivfpq(tables, nullptr, d_queries, num_queries, nullptr, top_k, nullptr, np,
                         do_rerank, (int)metric, d_valid_mask, d_out_results, top_k, ws.ptr, p, s);
    return e == hipSuccess ? 0 : (int)e;
}

int fv_ivf_pq_list_order_codes(const IvfPqGpuTables* tables, uint8_t* d_codes_out, void* stream) {
    if (!tables || !d_codes_out || !tables->d_pq_codes || !tables->d_ivf_vec_ids) return -1;
    size_t total = (size_t)tables->num_vectors * tables->m_pq;
    if (total == 0) return 0;
    fv::list_order_codes_kernel<<<(unsigned)((total + 255) / 256), 256, 0,
                                  static_cast<hipStream_t>(stream)>>>(
        tables->d_pq_codes, tables->d_ivf_vec_ids, tables->num_vectors, tables->m_pq, d_codes_out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

size_t fv_ivf_pq_workspace_size(const IvfPqGpuTables* tables, uint32_t num_queries,
                                uint32_t max_nprobe, uint32_t shortlist, uint32_t max_top_k) {
    if (!tables || max_top_k == 0 || max_top_k > FV_MAX_TOPK || shortlist > FV_MAX_SHORTLIST ||
        max_nprobe == 0 || tables->nlist == 0)
        return 0;
    uint32_t np = max_nprobe > tables->nlist ? tables->nlist : max_nprobe;
    return fv::plan_ivfpq(tables, num_queries, np, shortlist, max_top_k).workspace_bytes;
}

int fv_ivf_pq_search_tuned(const IvfPqGpuTables* tables, const uint8_t* d_codes_list_order,
                           const float* d_queries, uint32_t num_queries, const uint32_t* d_top_k,
                           uint32_t max_top_k, const uint32_t* d_nprobe, uint32_t max_nprobe,
                           uint32_t shortlist, int rerank, MetricType metric,
                           const uint32_t* d_valid_mask, GpuSearchResult* d_out_results,
                           void* d_workspace, size_t workspace_bytes,
                           const FvIvfPqTuning* tuning, void* stream) {
    if (!tables || !d_queries || !d_out_results) return -1;
    if (!fv::tables_ok(tables) || max_top_k == 0 || max_top_k > FV_MAX_TOPK ||
        shortlist > FV_MAX_SHORTLIST || max_nprobe == 0)
        return (int)hipErrorInvalidValue;
    if (num_queries == 0) return 0;
    uint32_t np = max_nprobe > tables->nlist ? tables->nlist : max_nprobe;
    fv::IvfPqPlan p = fv::plan_ivfpq(tables, num_queries, np, shortlist, max_top_k);
    if (!d_workspace || workspace_bytes < p.workspace_bytes) return (int)hipErrorInvalidValue;
    if (tuning) {
        p.target_items = tuning->target_items;
        if (tuning->chunk_min) p.chunk_min = tuning->chunk_min;
    }
    const int do_rerank = rerank && tables->d_vectors != nullptr;
    hipError_t e = fv::launch_ivfpq(tables, d_codes_list_order, d_queries, num_queries, d_top_k,
                                    max_top_k, d_nprobe, np, do_rerank, (int)metric, d_valid_mask,
                                    d_out_results, max_top_k, d_workspace, p,
                                    static_cast<hipStream_t>(stream));
    return e == hipSuccess ? 0 : (int)e;
}

int fv_ivf_pq_search(const IvfPqGpuTables* tables, const uint8_t* d_codes_list_order,
                     const float* d_queries, uint32_t num_queries, const uint32_t* d_top_k,
                     uint32_t max_top_k, const uint32_t* d_nprobe, uint32_t max_nprobe,
                     uint32_t shortlist, int rerank, MetricType metric,
                     const uint32_t* d_valid_mask, GpuSearchResult* d_out_results,
                     void* d_workspace, size_t workspace_bytes, void* stream) {
    return fv_ivf_pq_search_tuned(tables, d_codes_list_order, d_queries, num_queries, d_top_k,
                                  max_top_k, d_nprobe, max_nprobe, shortlist, rerank, metric,
                                  d_valid_mask, d_out_results, d_workspace, workspace_bytes,
                                  nullptr, stream);
}

}  // extern "C"
