// Reduced-precision flat search on the matrix cores (cuda_fp16_tc_search_batch
// and fv_flat_search): plan, query encoding, optional mask compaction, the
// scan launch (kernel in fv_flat_mfma_scan.hpp), and the select/rerank that
// turns the scan's per-chunk candidates into the final top-k.
//
// Reference: kernels/src/ivf_quant_lookup.cu:887-1058, 1106-1123.
//
//  * flat_select_kernel: per query, the candidates of every chunk that are
//    within the final shared bound are streamed into an LDS top-S.
//  * flat_rerank_kernel: per (query, 256 selected rows), exact FP32 distances
//    in the reference CPU path's order, rows staged through LDS.
//  * flat_finalize_kernel: per query, sort by (distance, row) and write the
//    first k. With FV_SEARCH_F64_BOUNDARY, candidates whose FP32 distance
//    lies within twice the FP32 rounding-error bound of a neighbour in that
//    order are re-scored in FP64 and reordered by (FP64 distance, row), so the
//    order of the returned rows, and which rows make the top k, follow the
//    exact distances.
//  * flat_tail_kernel: the three steps above fused into one wavefront per
//    query, used when the shortlist has at most 256 rows.
#include "fv_flat_mfma_scan.hpp"
#include "fv_workspace.hpp"

namespace fv {

// Threads of flat_select_kernel, one candidate each per round: 1024 for deep
// shortlists (many candidates per query), 256 otherwise.
constexpr uint32_t kSelectBatchSmall = 256;
constexpr uint32_t kSelectBatchLarge = 1024;

// Per query (one wavefront each, 4 per workgroup): a value with at least S of
// the first pass's subset minima at or below it (the S-th smallest, or
// slightly above it), found by bisection over the ordered distance with
// ballot counts; the minima stay in registers (NR per lane, m <= 64*NR) or,
// for NR = 0, in LDS (one wavefront per workgroup). Each minimum is the score
// of a distinct row, so at least S rows score at or below the result: it
// bounds the global S-th best and seeds the shared bound of the main pass.
template <int NR>
__global__ __launch_bounds__(256) void bound_select_kernel(const float* __restrict__ minima,
                                                           uint32_t m, uint32_t nq,
                                                           FlatSearchArgs a,
                                                           uint32_t* __restrict__ gthr) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t q = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (q >= nq) return;
    uint32_t k, S;
    query_k_s(a, q, k, S);
    constexpr int kR = NR > 0 ? NR : 1;
    uint32_t o[kR];
    extern __shared__ uint32_t s_o[];
    const uint32_t nr = (m + kWave - 1) / kWave;
    auto load = [&](uint32_t r) -> uint32_t {
        const uint32_t i = r * kWave + lane;
        uint32_t v = 0xFFFFFFFFu;
        if (i < m) {
            const float f = minima[((size_t)(i >> 4) * nq + q) * 16 + (i & 15u)];
            if (f < kAdmit) v = ord_f32(f);
        }
        return v;
    };
    uint32_t valid = 0;
    if (NR > 0) {
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            o[r] = load((uint32_t)r);
            valid += (uint32_t)__popcll(__ballot(o[r] != 0xFFFFFFFFu));
        }
    } else {
        for (uint32_t r = 0; r < nr; ++r) {
            const uint32_t v = load(r);
            s_o[r * kWave + lane] = v;
            valid += (uint32_t)__popcll(__ballot(v != 0xFFFFFFFFu));
        }
    }
    if (valid < S) return;
    uint32_t lo = 0u, hi = 0xFFFFFFFEu;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        uint32_t c = 0;
        if (NR > 0) {
#pragma unroll
            for (int r = 0; r < kR; ++r) c += (uint32_t)__popcll(__ballot(o[r] <= mid));
        } else {
            for (uint32_t r = 0; r < nr; ++r) c += (uint32_t)__popcll(__ballot(s_o[r * kWave + lane] <= mid));
        }
        if (c >= S) {
            hi = mid;
            if (c <= S + (S >> 3) + 16u) break;
        } else {
            lo = mid + 1u;
        }
    }
    if (lane == 0) atomicMin(&gthr[q], hi);
}

// Keys of rows whose distance is not admitted (>= 1e30 or NaN): the row in
// the low word keeps them distinct, and they order after every admitted key.
__device__ __forceinline__ uint64_t unadmitted_key(uint32_t row) { return 0xFFFFFFFF00000000ull | row; }
__device__ __forceinline__ bool key_unadmitted(uint64_t k) { return (uint32_t)(k >> 32) == 0xFFFFFFFFu; }

// Rounding of one query element as the scan's query encoder does it
// (encode_queries_block), decoded back to FP32.
__device__ __forceinline__ float coded_query_value(float v, uint32_t dt) {
    if (dt == FV_DTYPE_F16) return (float)(_Float16)v;
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}

// Bounded rerank set for the select path: 16-bit codes, L2 or inner product.
struct SelectBound {
    const float* qm;        // FP32 queries [nq][dim]
    const float* norms;     // |x|^2 per row
    const float* enc_err;   // |x~ - x| per row, or null (element bound)
    uint32_t dim;
    uint32_t code_dtype;    // FV_DTYPE_F16 / FV_DTYPE_BF16; 0 = no bound
    int metric;
};
constexpr int kSelectBoundPerThread = 8;

// One workgroup per query: streams the candidates of every chunk that are
// within the final shared bound into an LDS top-S (sorted only when more
// than S arrive) and writes the S best keys (approximate score, row) to
// sel[q][0..have) and have to sel_n[q]. With a bound (sb.code_dtype != 0)
// only the rows whose error interval reaches the k-th smallest upper end are
// written (the bound of bounded_rerank_set, applied by the whole workgroup):
// the others cannot be among the k best.
template <uint32_t kSelectBatch>
__global__ __launch_bounds__(kSelectBatch) void flat_select_kernel(
    const uint64_t* __restrict__ cand, const uint32_t* __restrict__ counts, uint32_t n_chunks,
    uint32_t nq, uint32_t cap, const uint32_t* __restrict__ gthr, FlatSearchArgs a,
    uint32_t lcap, uint64_t* __restrict__ sel, uint32_t* __restrict__ sel_n, SelectBound sb) {
    extern __shared__ uint64_t s_k[];
    __shared__ uint32_t s_n;
    __shared__ uint32_t s_g;
    __shared__ BlockSelectLds s_bs;
    const uint32_t q = blockIdx.x;
    uint32_t k, S;
    query_k_s(a, q, k, S);
    if (threadIdx.x == 0) {
        s_n = 0;
        s_g = gthr[q];
    }
    __syncthreads();
    const uint64_t cstride = a.cand_off ? a.cand_stride : (uint64_t)nq * cap;
    const uint64_t qbase = a.cand_off ? a.cand_off[q] : (uint64_t)q * cap;
    // Candidate counts of every chunk, prefix-summed in LDS (chunks processed
    // kSelectBatch at a time), then the concatenated candidates streamed in
    // rounds of kSelectBatch: one barrier per round instead of one per chunk.
    __shared__ uint32_t s_pre[kSelectBatch + 1];
    for (uint32_t c0 = 0; c0 < n_chunks; c0 += kSelectBatch) {
        const uint32_t nc = min(kSelectBatch, n_chunks - c0);
        const uint32_t my = threadIdx.x < nc ? counts[(size_t)(c0 + threadIdx.x) * nq + q] : 0u;
        s_pre[threadIdx.x + 1] = my;
        if (threadIdx.x == 0) s_pre[0] = 0;
        __syncthreads();
        for (uint32_t off = 1; off < kSelectBatch; off <<= 1) {
            const uint32_t v = threadIdx.x + 1 > off ? s_pre[threadIdx.x + 1 - off] : 0u;
            __syncthreads();
            s_pre[threadIdx.x + 1] += v;
            __syncthreads();
        }
        const uint32_t total = s_pre[nc];
        for (uint32_t e0 = 0; e0 < total; e0 += kSelectBatch) {
            const uint32_t e = e0 + threadIdx.x;
            const uint32_t g = s_g;
            if (e < total) {
                // Chunk of entry e: the last c with s_pre[c] <= e.
                uint32_t lo = 0, hi = nc - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (s_pre[mid] <= e) lo = mid;
                    else hi = mid - 1;
                }
                const uint64_t key = cand[(c0 + lo) * cstride + qbase + (e - s_pre[lo])];
                if ((uint32_t)(key >> 32) <= g) s_k[atomicAdd(&s_n, 1u)] = key;
            }
            __syncthreads();
            if (s_n + kSelectBatch > lcap) {
                // Cut to the S smallest (radix selection, no sort) and lower
                // the admission bound to the S-th smallest's distance.
                const uint32_t n = s_n;
                if (n > S) {
                    const uint64_t t = block_select(s_k, n, S, s_bs);
                    __syncthreads();
                    if (threadIdx.x == 0) {
                        s_n = S;
                        s_g = min(s_g, (uint32_t)(t >> 32));
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
    // The S smallest, in no particular order: the rerank re-scores them all
    // and the finalize orders them.
    const uint32_t n = s_n;
    if (n > S) {
        block_select(s_k, n, S, s_bs);
        __syncthreads();
    }
    const uint32_t have = n < S ? n : S;
    uint64_t* o = sel + (size_t)q * a.s_max;
    if (sb.code_dtype == 0 || have <= k || have > kSelectBoundPerThread * kSelectBatch) {
        for (uint32_t i = threadIdx.x; i < have; i += blockDim.x) o[i] = s_k[i];
        if (threadIdx.x == 0) sel_n[q] = have;
        return;
    }
    // The query's norm and coding error (FP64 sums; order does not matter,
    // only a bound is needed).
    constexpr uint32_t kWaves = kSelectBatch / kWave;
    __shared__ double s_red[2][kWaves];
    const float* qv = sb.qm + (size_t)q * sb.dim;
    double nq2 = 0.0, eq2 = 0.0;
    for (uint32_t d = threadIdx.x; d < sb.dim; d += kSelectBatch) {
        const float v = qv[d];
        const double e = (double)coded_query_value(-v, sb.code_dtype) + (double)v;
        nq2 = __builtin_fma((double)v, (double)v, nq2);
        eq2 = __builtin_fma(e, e, eq2);
    }
    for (int off = 32; off > 0; off >>= 1) {
        nq2 += __shfl_xor(nq2, off);
        eq2 += __shfl_xor(eq2, off);
    }
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    if (lane == 0) {
        s_red[0][wave] = nq2;
        s_red[1][wave] = eq2;
    }
    __syncthreads();
    nq2 = 0.0;
    eq2 = 0.0;
    for (uint32_t w = 0; w < kWaves; ++w) {
        nq2 += s_red[0][w];
        eq2 += s_red[1][w];
    }
    const double slack = 1.0 + 0x1p-40;
    const double qn = sqrt(nq2) * slack, eq = sqrt(eq2) * slack, qtn = qn + eq;
    const double g8 = 8.0 * (1.25 * (double)sb.dim + 8.0) * 0x1p-24;
    const double cm = sb.metric == METRIC_L2 ? 2.0 : 1.0;
    const double u_rel = (sb.code_dtype == FV_DTYPE_F16 ? 0x1p-11 : 0x1p-8) * (1.0 + 0x1p-10);
    const double u_abs = sb.code_dtype == FV_DTYPE_F16 ? 0x1p-24 * sqrt((double)sb.dim) * (1.0 + 0x1p-10) : 0.0;
    constexpr int NR = kSelectBoundPerThread;
    uint32_t hi_o[NR], lo_o[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const uint32_t i = (uint32_t)r * kSelectBatch + threadIdx.x;
        hi_o[r] = 0xFFFFFFFFu;  // past the list: never counted, never kept
        lo_o[r] = 0xFFFFFFFFu;
        if (i < have) {
            const uint64_t kv = s_k[i];
            const uint32_t row = key_slot(kv);
            const double sc = (double)key_dist(kv);
            const double E = sb.metric == METRIC_L2 ? nq2 + 2.0 * sc : sc;
            const double xn = sqrt((double)sb.norms[row]) * slack;
            const double ex = sb.enc_err ? (double)sb.enc_err[row] : u_rel * xn + u_abs;
            const double xtn = xn + ex;
            const double B = cm * (qtn * ex + eq * xn) + g8 * (qtn + xtn) * (qtn + xtn);
            const double hi = E + B, lo = E - B;
            if (key_unadmitted(kv) || !(hi < 1e300) || !(lo > -1e300) || !(B < 1e300)) {
                hi_o[r] = 0xFFFFFFFEu;  // counted as unbounded, always kept
                lo_o[r] = 0u;
            } else {
                float hf = (float)hi, lf = (float)lo;
                if ((double)hf < hi) hf = next_up(hf);
                if ((double)lf > lo) lf = -next_up(-lf);
                hi_o[r] = ord_f32(hf);
                lo_o[r] = ord_f32(lf);
            }
        }
    }
    // U: the smallest ordered value with at least k upper ends at or below it.
    uint32_t blo = 0u, bhi = 0xFFFFFFFEu;
    while (blo < bhi) {
        const uint32_t mid = blo + ((bhi - blo) >> 1);
        uint32_t c = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if ((uint32_t)r * kSelectBatch < have) c += (uint32_t)__syncthreads_count(hi_o[r] <= mid);
        if (c >= k) bhi = mid;
        else blo = mid + 1u;
    }
    __shared__ uint32_t s_m;
    if (threadIdx.x == 0) s_m = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const uint32_t i = (uint32_t)r * kSelectBatch + threadIdx.x;
        if (i < have && lo_o[r] <= bhi) o[atomicAdd(&s_m, 1u)] = s_k[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) sel_n[q] = s_m;
}

// Exact FP32 distance of one (query, row) pair, summed in the reference CPU
// path's order (sequential over d, non-fused: the library is built with
// -ffp-contract=off, so it equals exact_distance<METRIC> bit for bit), plus
// the running error bound of each recursive sum (Higham, Accuracy and
// Stability of Numerical Algorithms, 2002, sec. 4.2): |error| <= u * sum_i
// |partial sum_i| + c * u * sum_i |term_i| to first order, c = 1 for products
// and 3 for squared differences.
template <int METRIC>
struct RerankSum {
    float acc = 0.0f, nqa = 0.0f, nxa = 0.0f, sabs = 0.0f;
    float run = 0.0f, run_q = 0.0f, run_x = 0.0f;
    __device__ __forceinline__ void add(float a, float b) {
        if (METRIC == METRIC_L2) {
            const float diff = a - b;
            const float sq = diff * diff;
            acc = acc + sq;
            run += acc;
        } else {
            const float p = a * b;
            acc = acc + p;
            run += fabsf(acc);
            sabs += fabsf(p);
            if (METRIC == METRIC_COSINE) {
                const float pa = a * a, pb = b * b;
                nqa = nqa + pa;
                nxa = nxa + pb;
                run_q += nqa;
                run_x += nxa;
            }
        }
    }
    // Distance and its error bound.
    __device__ __forceinline__ void finish(float& d32, float& e) const {
        constexpr float u = 5.9604645e-8f;  // 2^-24
        constexpr float margin = 1.01f;     // covers the second-order terms
        if (METRIC == METRIC_L2) {
            d32 = acc;
            e = margin * u * (run + 3.0f * acc);
        } else if (METRIC == METRIC_INNER_PRODUCT) {
            d32 = -acc;
            e = margin * u * (run + sabs);
        } else {
            const float denom = sqrtf(nqa) * sqrtf(nxa) + 1e-8f;
            d32 = 1.0f - acc / denom;
            const float e_dot = u * (run + sabs);
            // Relative error of the denominator: half of each norm's relative
            // error, plus two square roots, a product and an add.
            const float rel = 0.5f * u * ((run_q + nqa) / fmaxf(nqa, 1e-30f) +
                                          (run_x + nxa) / fmaxf(nxa, 1e-30f)) + 4.0f * u;
            e = margin * ((e_dot + fabsf(acc) * rel) / denom + u * fabsf(acc / denom) + u * fabsf(d32));
        }
    }
};

// Distance reported for a selected row when no FP32 rows are stored: the
// reduced-precision score s (0.5*||x||^2 - q.x for L2, -q.x for IP,
// -q.x/||x|| for cosine) turned into a distance.
template <int METRIC>
__device__ __forceinline__ float score_as_distance(float s, const float* __restrict__ qv, uint32_t dim) {
    if (METRIC == METRIC_INNER_PRODUCT) return s;
    float qn = 0.0f;
    for (uint32_t dd = 0; dd < dim; ++dd) qn = __builtin_fmaf(qv[dd], qv[dd], qn);
    return METRIC == METRIC_L2 ? fmaxf(0.0f, qn + 2.0f * s) : 1.0f + s / (sqrtf(qn) + 1e-8f);
}

// Exact FP32 rerank of the selected rows. Workgroup b covers items
// [256 b, 256 b + 256) of the padded [query][s_max] array of selected keys,
// one item per thread (a workgroup may span several queries; items past a
// query's count are idle). Rows are staged through LDS in slabs of 32 dims
// with coalesced loads, the workgroup's queries too when they fit; each
// thread keeps its row's sums in the reference CPU path's order (sequential
// over d, non-fused), so the distance equals exact_distance<METRIC> bit for
// bit. Writes (distance, row) keys in place and, when err != NULL, the FP32
// error bound of each distance. Without FP32 rows the approximate score is
// turned into a distance instead.
constexpr uint32_t kRerankRows = 256;
constexpr uint32_t kRerankSlab = 32;
constexpr uint32_t kRerankQFloats = 4096;

template <int METRIC>
__global__ __launch_bounds__(256) void flat_rerank_kernel(
    const float* __restrict__ X32, const float* __restrict__ Qm, uint32_t dim, uint32_t nq,
    uint32_t smax, const uint32_t* __restrict__ sel_n, const uint32_t* __restrict__ sel_off,
    const uint32_t* __restrict__ sel_total, uint64_t* __restrict__ sel, float* __restrict__ err) {
    __shared__ float s_x[kRerankRows * (kRerankSlab + 1)];
    __shared__ float s_qv[kRerankQFloats];
    __shared__ uint32_t s_row[kRerankRows];
    __shared__ uint8_t s_act[kRerankRows];
    // Items are the selected rows of all queries, concatenated (sel_off is
    // the exclusive prefix of the per-query counts): every workgroup but the
    // last gets 256 rows.
    const uint32_t tid = threadIdx.x;
    const uint32_t total = *sel_total;
    const uint32_t first = blockIdx.x * kRerankRows;
    if (first >= total) return;
    const uint32_t it = first + tid;
    const bool act = it < total;
    // Query of item `it`: the last q with sel_off[q] <= it.
    uint32_t q = 0;
    {
        uint32_t lo = 0, hi = nq - 1;
        const uint32_t key = act ? it : total - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (sel_off[mid] <= key) lo = mid;
            else hi = mid - 1;
        }
        q = lo;
    }
    const size_t item = (size_t)q * smax + (act ? it - sel_off[q] : 0u);
    const uint64_t my_key = act ? sel[item] : kEmptyKey;
    s_row[tid] = act ? key_slot(my_key) : 0u;
    s_act[tid] = act ? 1 : 0;
    // Queries of this workgroup: q0 .. q1 (items are in query order).
    __shared__ uint32_t s_q01[2];
    if (tid == 0) s_q01[0] = q;
    if (it == min(first + kRerankRows, total) - 1) s_q01[1] = q;
    __syncthreads();
    const uint32_t q0 = s_q01[0], q1 = s_q01[1];
    const bool staged = (size_t)(q1 - q0 + 1) * dim <= kRerankQFloats;
    if (staged)
        for (uint32_t i = tid; i < (q1 - q0 + 1) * dim; i += blockDim.x) s_qv[i] = Qm[(size_t)q0 * dim + i];
    __syncthreads();
    const float* qv = act ? (staged ? &s_qv[(q - q0) * dim] : Qm + (size_t)q * dim) : s_qv;
    float d32 = 0.0f;
    float e = 0.0f;
    if (X32) {
        RerankSum<METRIC> sum;
        const bool vec = (dim & 3u) == 0u && (reinterpret_cast<uintptr_t>(X32) & 15u) == 0u;
        for (uint32_t d0 = 0; d0 < dim; d0 += kRerankSlab) {
            const uint32_t w = min(kRerankSlab, dim - d0);
            // 256 rows x 32 dims: 8 float4 (or 32 scalars) per thread.
            if (vec) {
#pragma unroll
                for (uint32_t m = 0; m < 8; ++m) {
                    const uint32_t jj = tid + 256u * m;
                    const uint32_t row = jj >> 3, part = jj & 7u;
                    if (s_act[row] && part * 4u < w) {
                        const float4 v = *reinterpret_cast<const float4*>(
                            X32 + (size_t)s_row[row] * dim + d0 + part * 4u);
                        float* dst = &s_x[row * (kRerankSlab + 1) + part * 4u];
                        dst[0] = v.x;
                        dst[1] = v.y;
                        dst[2] = v.z;
                        dst[3] = v.w;
                    }
                }
            } else {
                for (uint32_t m = 0; m < kRerankSlab; ++m) {
                    const uint32_t jj = tid + 256u * m;
                    const uint32_t row = jj / kRerankSlab, c = jj % kRerankSlab;
                    if (s_act[row] && c < w)
                        s_x[row * (kRerankSlab + 1) + c] = X32[(size_t)s_row[row] * dim + d0 + c];
                }
            }
            __syncthreads();
            if (act) {
                const float* xr = &s_x[tid * (kRerankSlab + 1)];
                for (uint32_t dd = 0; dd < w; ++dd) sum.add(qv[d0 + dd], xr[dd]);
            }
            __syncthreads();
        }
        sum.finish(d32, e);
    } else if (act) {
        // No FP32 rows: report the reduced-precision score as a distance.
        d32 = score_as_distance<METRIC>(key_dist(my_key), qv, dim);
    }
    if (act) {
        sel[item] = d32 < kAdmit ? make_key(d32, s_row[tid]) : unadmitted_key(s_row[tid]);
        if (err) err[item] = e;
    }
}

// One workgroup per query: sorts the reranked keys by (distance, row) and
// writes the first k. With FV_SEARCH_F64_BOUNDARY, two candidates adjacent
// in FP32 order whose gap is at most twice the largest error bound may be in
// either exact order: chains of such gaps form clusters, which are re-scored
// in FP64 and reordered by (FP64 distance, row). Outside a cluster the FP32
// order is the exact order, so only clusters that start before position k are
// touched.
// Up to 16 wavefronts: with deep shortlists the kernel's LDS allows one
// workgroup per CU, and the FP64 re-scoring's row reads need many waves in
// flight; shallow shortlists run 4 wavefronts per workgroup (the kernel
// strides by blockDim.x).
constexpr int kFinalThreads = 1024;
// Deepest shortlist finalized by a 256-thread workgroup when the finalize is
// split by depth (a power of two: it is that launch's key capacity).
constexpr uint32_t kFinalSplitS = 512;
template <int METRIC>
__global__ __launch_bounds__(kFinalThreads) void flat_finalize_kernel(
    const uint64_t* __restrict__ sel, const uint32_t* __restrict__ sel_n,
    const float* __restrict__ err, uint32_t hcap, FlatSearchArgs a,
    const float* __restrict__ X32, const float* __restrict__ Qm,
    GpuSearchResult* __restrict__ out, uint32_t s_lo, uint32_t s_hi) {
    extern __shared__ uint64_t s_k[];
    __shared__ uint32_t s_emax;
    __shared__ BlockSelectLds s_bs;
    __shared__ uint64_t s_succ[kSelMaxWaves];
    const uint32_t q = blockIdx.x;
    uint32_t k, S;
    query_k_s(a, q, k, S);
    // Queries whose shortlist depth is outside [s_lo, s_hi) belong to the
    // other launch (the whole workgroup leaves before any barrier).
    if (S < s_lo || S >= s_hi) return;
    const uint32_t have_in = sel_n[q];
    const uint64_t* src = sel + (size_t)q * a.s_max;
    for (uint32_t i = threadIdx.x; i < have_in; i += blockDim.x) s_k[i] = src[i];
    const bool f64b = (a.flags & FV_SEARCH_F64_BOUNDARY) && X32 != nullptr && err != nullptr;
    if (threadIdx.x == 0) s_emax = 0u;
    __syncthreads();
    if (f64b) {
        // Largest FP32 error bound (positive floats order like their bits).
        float e = 0.0f;
        for (uint32_t i = threadIdx.x; i < have_in; i += blockDim.x)
            e = fmaxf(e, err[(size_t)q * a.s_max + i]);
        atomicMax(&s_emax, __float_as_uint(e));
    }
    __syncthreads();
    // Only the first k in (distance, row) order are written. Without the FP64
    // boundary the k smallest are selected and only they are sorted. With it,
    // every candidate within band of the k-th (all that can be re-scored) is
    // kept, plus the next one above the band (the neighbour the band test
    // looks at); a candidate further up cannot reach the first k, whose
    // values (FP32, or FP64 within the error bound) stay below the band.
    const uint32_t kq = k < have_in ? k : have_in;
    uint32_t have = have_in;
    if (have_in > kq && kq > 0) {
        if (!(f64b && have_in > 1)) {
            block_select(s_k, have_in, kq, s_bs);
            have = kq;
        } else {
            uint64_t pmask, prefix;
            block_radix_kth(s_k, have_in, kq, s_bs, pmask, prefix);
            for (uint32_t i = threadIdx.x; i < have_in; i += blockDim.x)
                if ((s_k[i] & pmask) == prefix) s_bs.thr = s_k[i];
            __syncthreads();
          if (!key_unadmitted(s_bs.thr)) {   // else fewer than k admitted: sort them all
            const float lim = key_dist(s_bs.thr) + 2.0000002f * __uint_as_float(s_emax);
            unsigned long long succ = kEmptyKey;
            for (uint32_t i = threadIdx.x; i < have_in; i += blockDim.x) {
                const uint64_t key = s_k[i];
                if (key_dist(key) > lim && key < succ) succ = key;
            }
#pragma unroll
            for (int off = kWave / 2; off > 0; off >>= 1) {
                const unsigned long long o = __shfl_xor(succ, off);
                succ = o < succ ? o : succ;
            }
            if ((threadIdx.x & (kWave - 1)) == 0) s_succ[threadIdx.x / kWave] = succ;
            __syncthreads();
            succ = s_succ[0];
            for (uint32_t w = 1; w < blockDim.x / kWave; ++w) succ = s_succ[w] < succ ? s_succ[w] : succ;
            const uint32_t m = block_compact(s_k, have_in, [lim](uint64_t key) {
                return key_dist(key) <= lim ? 1 : 0;
            }, s_bs);
            __syncthreads();
            if (succ != kEmptyKey && threadIdx.x == 0) s_k[m] = succ;
            have = m + (succ != kEmptyKey ? 1u : 0u);
          }
        }
        __syncthreads();
    }
    const uint32_t hp = next_pow2(have < 2u ? 2u : have);
    for (uint32_t i = have + threadIdx.x; i < hp; i += blockDim.x) s_k[i] = kEmptyKey;
    __syncthreads();
    if (have <= blockDim.x) {
        // Each key to its rank (keys are distinct).
        const uint64_t mine = threadIdx.x < have ? s_k[threadIdx.x] : kEmptyKey;
        uint32_t rank = 0;
        if (threadIdx.x < have)
            for (uint32_t j = 0; j < have; ++j) rank += s_k[j] < mine ? 1u : 0u;
        __syncthreads();
        if (threadIdx.x < have) s_k[rank] = mine;
        __syncthreads();
    } else {
        block_bitonic_sort(s_k, hp);
    }
    const float* qv = Qm + (size_t)q * a.dim;
    if (f64b && have > 1) {
        // Candidates adjacent in FP32 order whose gap is at most twice the
        // largest error bound may be in either exact order. Every such
        // candidate that could reach the top k (FP32 distance at most the
        // k-th plus that band) is re-scored in FP64; then all candidates are
        // sorted by (value, row) with value = the FP64 distance where
        // re-scored, else the FP32 distance. A candidate that is not
        // re-scored is more than the band away from its neighbours, so its
        // place in this order is its exact place.
        uint64_t* s_v = s_k + hcap;                               // ordered value bits
        uint32_t* s_r = reinterpret_cast<uint32_t*>(s_v + hcap);  // rows
        __shared__ uint32_t s_any;
        const float band = 2.0000002f * __uint_as_float(s_emax);
        const uint32_t kk = k < have ? k : have;
        const float lim = key_dist(s_k[kk - 1]) + band;
        if (threadIdx.x == 0) s_any = 0u;
        __syncthreads();
        // Pass 1: which candidates are re-scored (s_r bit 31 marks them; rows
        // are < 2^31 here, checked by the caller).
        for (uint32_t i = threadIdx.x; i < hp; i += blockDim.x) {
            const uint64_t key = s_k[i];
            uint32_t r = 0xFFFFFFFFu;
            uint64_t v = ~0ull;
            if (i < have && !key_unadmitted(key)) {
                const float d = key_dist(key);
                r = key_slot(key);
                const bool near_prev = i > 0 && !key_unadmitted(s_k[i - 1]) && d - key_dist(s_k[i - 1]) <= band;
                const bool near_nxt = i + 1 < have && !key_unadmitted(s_k[i + 1]) && key_dist(s_k[i + 1]) - d <= band;
                const uint64_t bits = __double_as_longlong((double)d);
                v = (bits & 0x8000000000000000ull) ? ~bits : (bits | 0x8000000000000000ull);
                if ((near_prev || near_nxt) && d <= lim) {
                    r |= 0x80000000u;
                    s_any = 1u;
                }
            }
            s_v[i] = v;
            s_r[i] = r;
        }
        __syncthreads();
        // Pass 2: FP64 distances of the marked candidates. Wave w takes the
        // entries w, w + W, w + 2W, ... (W waves), up to 4 marked ones at a
        // time so that their row loads overlap; lanes split the dims and the
        // sums are reduced across the wave.
        if (s_any) {
            constexpr int kRows = 4;
            const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
            const uint32_t nw = blockDim.x / kWave;
            uint32_t i = wave;
            while (i < have) {
                uint32_t idx[kRows], row[kRows];
                int m = 0;
                for (; i < have && m < kRows; i += nw) {
                    const uint32_t r = s_r[i];
                    if (r == 0xFFFFFFFFu || !(r & 0x80000000u)) continue;
#pragma unroll
                    for (int t = 0; t < kRows; ++t)
                        if (t == m) {
                            idx[t] = i;
                            row[t] = r & 0x7FFFFFFFu;
                        }
                    ++m;
                }
                double acc[kRows], nqd[kRows], nxd[kRows];
#pragma unroll
                for (int t = 0; t < kRows; ++t) acc[t] = nqd[t] = nxd[t] = 0.0;
                // Each lane sums its dims d = lane, lane + 64, ... in order;
                // 4 of them are loaded per step so that the loads overlap.
                auto fma_step = [&](double qd, const double (&xd)[kRows]) {
#pragma unroll
                    for (int t = 0; t < kRows; ++t) {
                        if (METRIC == METRIC_L2) {
                            const double u = qd - xd[t];
                            acc[t] = __builtin_fma(u, u, acc[t]);
                        } else {
                            acc[t] = __builtin_fma(qd, xd[t], acc[t]);
                            if (METRIC == METRIC_COSINE) {
                                nqd[t] = __builtin_fma(qd, qd, nqd[t]);
                                nxd[t] = __builtin_fma(xd[t], xd[t], nxd[t]);
                            }
                        }
                    }
                };
                uint32_t d = lane;
                for (; d + 3u * kWave < a.dim; d += 4u * kWave) {
                    double qd[4], xd[4][kRows];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        qd[u] = qv[d + u * kWave];
#pragma unroll
                        for (int t = 0; t < kRows; ++t)
                            xd[u][t] = t < m ? (double)X32[(size_t)row[t] * a.dim + d + u * kWave] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) fma_step(qd[u], xd[u]);
                }
                for (; d < a.dim; d += kWave) {
                    double xd[kRows];
#pragma unroll
                    for (int t = 0; t < kRows; ++t)
                        xd[t] = t < m ? (double)X32[(size_t)row[t] * a.dim + d] : 0.0;
                    fma_step(qv[d], xd);
                }
#pragma unroll
                for (int t = 0; t < kRows; ++t) {
                    for (int off = 32; off > 0; off >>= 1) {
                        acc[t] += __shfl_xor(acc[t], off);
                        if (METRIC == METRIC_COSINE) {
                            nqd[t] += __shfl_xor(nqd[t], off);
                            nxd[t] += __shfl_xor(nxd[t], off);
                        }
                    }
                }
                if (lane == 0) {
#pragma unroll
                    for (int t = 0; t < kRows; ++t) {
                        if (t >= m) break;
                        double dd;
                        if (METRIC == METRIC_L2) dd = acc[t];
                        else if (METRIC == METRIC_INNER_PRODUCT) dd = -acc[t];
                        else dd = 1.0 - acc[t] / (sqrt(nqd[t]) * sqrt(nxd[t]) + 1e-8);
                        const uint64_t bits = __double_as_longlong(dd);
                        s_v[idx[t]] = (bits & 0x8000000000000000ull) ? ~bits : (bits | 0x8000000000000000ull);
                    }
                }
            }
        }
        __syncthreads();
        if (s_any) {
            // Re-scored candidates (bit 31 of s_r, kept until written out)
            // form runs of adjacent entries; an entry outside a run is more
            // than the band away from both neighbours and every re-scored
            // value is within the error bound of its FP32 value, so each run
            // only reorders among itself (entries above lim, which never
            // reach the first k, aside). Short runs are sorted by their first
            // thread; a run longer than kMaxRun sends the workgroup to a
            // full sort of (value, row) pairs.
            constexpr uint32_t kMaxRun = 32;
            __shared__ uint32_t s_maxrun;
            auto marked = [&](uint32_t i) {
                const uint32_t r = s_r[i];
                return r != 0xFFFFFFFFu && (r & 0x80000000u) != 0u;
            };
            if (threadIdx.x == 0) s_maxrun = 0u;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < have; i += blockDim.x) {
                if (!marked(i) || (i > 0 && marked(i - 1))) continue;
                uint32_t j = i + 1;
                while (j < have && marked(j)) ++j;
                atomicMax(&s_maxrun, j - i);
            }
            __syncthreads();
            if (s_maxrun <= kMaxRun) {
                for (uint32_t i = threadIdx.x; i < have; i += blockDim.x) {
                    if (!marked(i) || (i > 0 && marked(i - 1))) continue;
                    uint32_t j = i + 1;
                    while (j < have && marked(j)) ++j;
                    for (uint32_t a2 = i + 1; a2 < j; ++a2) {
                        const uint64_t v = s_v[a2];
                        const uint32_t r = s_r[a2];
                        uint32_t b2 = a2;
                        while (b2 > i && (s_v[b2 - 1] > v || (s_v[b2 - 1] == v &&
                                                              (s_r[b2 - 1] & 0x7FFFFFFFu) > (r & 0x7FFFFFFFu)))) {
                            s_v[b2] = s_v[b2 - 1];
                            s_r[b2] = s_r[b2 - 1];
                            --b2;
                        }
                        s_v[b2] = v;
                        s_r[b2] = r;
                    }
                }
            } else {
                for (uint32_t i = threadIdx.x; i < hp; i += blockDim.x)
                    if (s_r[i] != 0xFFFFFFFFu) s_r[i] &= 0x7FFFFFFFu;
                __syncthreads();
                block_bitonic_sort_pairs(s_v, s_r, hp);
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < kk; i += blockDim.x) {
                const uint64_t v = s_v[i];
                const uint64_t bits = (v & 0x8000000000000000ull) ? (v & 0x7fffffffffffffffull) : ~v;
                const uint32_t r = s_r[i] == 0xFFFFFFFFu ? 0xFFFFFFFFu : (s_r[i] & 0x7FFFFFFFu);
                s_k[i] = make_key((float)__longlong_as_double(bits), r);
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < a.top_k; i += blockDim.x) {
        const uint64_t key = i < k && i < hp ? s_k[i] : kEmptyKey;
        GpuSearchResult r;
        // Empty slots, rows whose distance is not admitted, and the FP64
        // path's placeholders (row 0xFFFFFFFF) are written empty.
        if (key_unadmitted(key) || key_slot(key) == FV_EMPTY_ID || !(key_dist(key) < kAdmit)) {
            r.id = FV_EMPTY_ID;
            r.distance = FV_EMPTY_DIST;
        } else {
            r.id = key_slot(key);
            r.distance = key_dist(key);
        }
        out[(size_t)q * a.top_k + i] = r;
    }
}

// Fused select + rerank + finalize for shortlists of at most kTailMaxS rows:
// one wavefront per query, kTailWaves per workgroup, no workgroup barriers.
// The wave streams the candidates of every chunk that are within the final
// shared bound into a wave-private LDS list kept to the S smallest keys by
// bitonic prunes, re-scores the selected rows exactly (lane i takes rows i,
// i+64, ...; two rows at a time so that their loads overlap), sorts them by
// (distance, row), applies the FP64 boundary when asked and writes the first
// k: the same results as flat_select_kernel, flat_rerank_kernel and
// flat_finalize_kernel in sequence.
constexpr uint32_t kTailMaxS = 256;
// Keys of a per-query candidate buffer beyond the query's shortlist: 64 of
// bisection slack plus 512 rows between buffer checks.
constexpr uint32_t kPerQuerySlack = 64u + 512u;
constexpr uint32_t kTailWaves = 4;

// Exact distances and error bounds of NR rows of one query (rows[] valid,
// possibly repeated), summed as RerankSum does. Rows with live[r] false are
// not read (their results are garbage and must be dropped): a load
// instruction then touches only the lines of the lanes that need a row.
template <int METRIC, int NR>
__device__ __forceinline__ void rerank_rows(const float* __restrict__ X32, const float* __restrict__ qv,
                                            uint32_t dim, bool vec, const uint32_t (&rows)[NR],
                                            const bool (&live)[NR], float (&d32)[NR],
                                            float (&err)[NR]) {
    RerankSum<METRIC> sum[NR];
    if (vec) {
        const float* xr[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) xr[r] = X32 + (size_t)rows[r] * dim;
        uint32_t d = 0;
        for (; d + 32u <= dim; d += 32u) {
            float4 q4[8], x4[NR][8];
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                q4[m] = *reinterpret_cast<const float4*>(qv + d + 4 * m);
#pragma unroll
                for (int r = 0; r < NR; ++r)
                    x4[r][m] = live[r] ? *reinterpret_cast<const float4*>(xr[r] + d + 4 * m)
                                       : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
#pragma unroll
            for (int m = 0; m < 8; ++m)
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    sum[r].add(q4[m].x, x4[r][m].x);
                    sum[r].add(q4[m].y, x4[r][m].y);
                    sum[r].add(q4[m].z, x4[r][m].z);
                    sum[r].add(q4[m].w, x4[r][m].w);
                }
        }
        for (; d < dim; d += 4u) {
            const float4 q4 = *reinterpret_cast<const float4*>(qv + d);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const float4 x4 = live[r] ? *reinterpret_cast<const float4*>(xr[r] + d)
                                          : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                sum[r].add(q4.x, x4.x);
                sum[r].add(q4.y, x4.y);
                sum[r].add(q4.z, x4.z);
                sum[r].add(q4.w, x4.w);
            }
        }
    } else {
        for (uint32_t d = 0; d < dim; ++d) {
            const float a = qv[d];
#pragma unroll
            for (int r = 0; r < NR; ++r) sum[r].add(a, live[r] ? X32[(size_t)rows[r] * dim + d] : 0.0f);
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) sum[r].finish(d32[r], err[r]);
}

// Ranks of the keys at positions lane, lane+64, ... (< n) among keys[0..n)
// of a wave-private LDS list of distinct keys: one broadcast read per key and
// one 64-bit compare per slot, no sorting network. Slots at or past n get
// rank n or more.
constexpr int kRankSlots = 4;
__device__ __forceinline__ void wave_key_ranks(const uint64_t* keys, uint32_t n, uint32_t lane,
                                               uint64_t (&mine)[kRankSlots], uint32_t (&rank)[kRankSlots]) {
    const uint32_t ns = (n + kWave - 1) / kWave;
#pragma unroll
    for (int s = 0; s < kRankSlots; ++s) {
        const uint32_t idx = lane + (uint32_t)s * kWave;
        mine[s] = idx < n ? keys[idx] : kEmptyKey;
        rank[s] = idx < n ? 0u : n;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t ki = keys[i];
#pragma unroll
        for (int s = 0; s < kRankSlots; ++s)
            if ((uint32_t)s < ns) rank[s] += ki < mine[s] ? 1u : 0u;
    }
}

// The kth smallest (1 <= kth <= count) of distinct keys held NR per lane
// (kEmptyKey past the list): bisection over the distance bits, then, when
// several keys share that distance, over the row bits among them.
template <int NR>
__device__ __forceinline__ uint64_t wave_kth_key(const uint64_t (&v)[NR], uint32_t kth) {
    uint32_t lo = 0u, hi = 0xFFFFFFFFu;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        uint32_t c = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) c += (uint32_t)__popcll(__ballot((uint32_t)(v[r] >> 32) <= mid));
        if (c >= kth) hi = mid;
        else lo = mid + 1u;
    }
    const uint32_t dk = hi;
    uint32_t cle = 0;
#pragma unroll
    for (int r = 0; r < NR; ++r) cle += (uint32_t)__popcll(__ballot((uint32_t)(v[r] >> 32) <= dk));
    if (cle == kth) return ((uint64_t)dk << 32) | 0xFFFFFFFFull;
    uint32_t slo = 0u, shi = 0xFFFFFFFFu;
    while (slo < shi) {
        const uint32_t mid = slo + ((shi - slo) >> 1);
        uint32_t c = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const uint32_t d = (uint32_t)(v[r] >> 32);
            c += (uint32_t)__popcll(__ballot(d < dk || (d == dk && (uint32_t)v[r] <= mid)));
        }
        if (c >= kth) shi = mid;
        else slo = mid + 1u;
    }
    return ((uint64_t)dk << 32) | shi;
}

// Bounded rerank set of one query (one wavefront; keys[0..have) hold the
// shortlist by scan score, have <= 256, k < have). For L2 the scan score is
// s = 0.5|x|^2 - q~.x~ accumulated in FP32 from the FP32 norm (q~, x~ the
// FP16/BF16 codes), so E = |q|^2 + 2s estimates the exact distance; for the
// inner product E = s = -q~.x~. The value R the exact FP32 rerank returns
// satisfies |R - E| <= B with
//   B = c_m (|q~| e_x + e_q |x|) + 8 g (|q~| + |x~|)^2,   g = (1.25 d + 8) 2^-24,
// c_m = 2 (L2) or 1 (IP): the first term bounds |q~.x~ - q.x| (e_x =
// |x~ - x| from enc_err or, without it, half an ulp per element, 2^-11 |x|
// (FP16; 2^-8 BF16) plus 2^-24 sqrt(d) for truncated FP16 subnormals: codes
// rounded to nearest, as the encoders here and the reference's encode_fp16_padded
// produce them; e_q = |q~ - q| computed here), the
// second, with a factor-of-8 margin, the FP32 rounding of the MFMA
// accumulation, of the norms and of the reference-order sums of the rerank
// (each at most g times (|q| + |x|)^2). With U the k-th smallest E + B, a row
// with E - B > U has k rows with R <= U < its own R, so it is not among the k
// best. Those rows get keys above every admitted one; the others are moved to
// keys[0..m) and m (>= k) is returned. Non-finite scores are always reranked.
template <int METRIC>
__device__ __noinline__ uint32_t bounded_rerank_set(uint64_t* keys, uint32_t have, uint32_t k,
                                                    const float* __restrict__ qv, uint32_t dim,
                                                    const float* __restrict__ norms,
                                                    const float* __restrict__ enc_err,
                                                    uint32_t code_dtype, uint32_t lane) {
    double nq2 = 0.0, eq2 = 0.0;
    for (uint32_t d = lane; d < dim; d += kWave) {
        const float v = qv[d];
        const double e = (double)coded_query_value(-v, code_dtype) + (double)v;
        nq2 = __builtin_fma((double)v, (double)v, nq2);
        eq2 = __builtin_fma(e, e, eq2);
    }
    for (int off = 32; off > 0; off >>= 1) {
        nq2 += __shfl_xor(nq2, off);
        eq2 += __shfl_xor(eq2, off);
    }
    const double slack = 1.0 + 0x1p-40;
    const double qn = sqrt(nq2) * slack, eq = sqrt(eq2) * slack, qtn = qn + eq;
    const double g8 = 8.0 * (1.25 * (double)dim + 8.0) * 0x1p-24;
    const double cm = METRIC == METRIC_L2 ? 2.0 : 1.0;
    const double u_rel = (code_dtype == FV_DTYPE_F16 ? 0x1p-11 : 0x1p-8) * (1.0 + 0x1p-10);
    const double u_abs = code_dtype == FV_DTYPE_F16 ? 0x1p-24 * sqrt((double)dim) * (1.0 + 0x1p-10) : 0.0;
    constexpr int NR = 4;  // have <= 256 = NR * 64
    uint64_t kv[NR];
    uint32_t hi_o[NR], lo_o[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const uint32_t i = (uint32_t)r * kWave + lane;
        kv[r] = i < have ? keys[i] : kEmptyKey;
        hi_o[r] = 0xFFFFFFFFu;  // past the list: never counted, never reranked
        lo_o[r] = 0xFFFFFFFFu;
        if (i < have) {
            const uint32_t row = key_slot(kv[r]);
            const double sc = (double)key_dist(kv[r]);
            const double E = METRIC == METRIC_L2 ? nq2 + 2.0 * sc : sc;
            const double xn = sqrt((double)norms[row]) * slack;
            const double ex = enc_err ? (double)enc_err[row] : u_rel * xn + u_abs;
            const double xtn = xn + ex;
            const double B = cm * (qtn * ex + eq * xn) + g8 * (qtn + xtn) * (qtn + xtn);
            const double hi = E + B, lo = E - B;
            if (key_unadmitted(kv[r]) || !(hi < 1e300) || !(lo > -1e300) || !(B < 1e300)) {
                hi_o[r] = 0xFFFFFFFEu;  // counted as unbounded, always reranked
                lo_o[r] = 0u;
            } else {
                // FP32 rounded outward, as ordered bits.
                float hf = (float)hi, lf = (float)lo;
                if ((double)hf < hi) hf = next_up(hf);
                if ((double)lf > lo) lf = -next_up(-lf);
                hi_o[r] = ord_f32(hf);
                lo_o[r] = ord_f32(lf);
            }
        }
    }
    // U: the smallest ordered value with at least k upper ends at or below it.
    uint32_t blo = 0u, bhi = 0xFFFFFFFEu;
    while (blo < bhi) {
        const uint32_t mid = blo + ((bhi - blo) >> 1);
        uint32_t c = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) c += (uint32_t)__popcll(__ballot(hi_o[r] <= mid));
        if (c >= k) bhi = mid;
        else blo = mid + 1u;
    }
    const uint32_t U = bhi;
    wave_lds_sync();
    // Rows to rerank first (in order), the others after them with keys above
    // every admitted key.
    uint32_t m = 0;
#pragma unroll
    for (int r = 0; r < NR; ++r) m += (uint32_t)__popcll(__ballot(lo_o[r] <= U));
    uint32_t pa = 0, pb = m;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const bool live = (uint32_t)r * kWave + lane < have;
        const bool re = live && lo_o[r] <= U;
        const bool out = live && !re;
        const uint64_t ba = __ballot(re), bb = __ballot(out);
        const uint32_t ra = __builtin_amdgcn_mbcnt_hi((uint32_t)(ba >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ba, 0u));
        const uint32_t rb = __builtin_amdgcn_mbcnt_hi((uint32_t)(bb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bb, 0u));
        if (re) keys[pa + ra] = kv[r];
        if (out) keys[pb + rb] = unadmitted_key(key_slot(kv[r]));
        pa += (uint32_t)__popcll(ba);
        pb += (uint32_t)__popcll(bb);
    }
    wave_lds_sync();
    return m;
}

template <int METRIC>
__global__ __launch_bounds__(256) void flat_tail_kernel(
    const uint64_t* __restrict__ cand, const uint32_t* __restrict__ counts, uint32_t n_chunks,
    uint32_t nq, uint32_t cap, const uint32_t* __restrict__ gthr, FlatSearchArgs a, uint32_t lw,
    uint32_t hw, uint32_t wbytes, const float* __restrict__ X32, const float* __restrict__ Qm,
    GpuSearchResult* __restrict__ out, const float* __restrict__ norms,
    const float* __restrict__ enc_err, uint32_t code_dtype) {
    extern __shared__ uint64_t s_mem[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    // Wave-uniform, so the query row is read with scalar loads.
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const uint32_t q = blockIdx.x * kTailWaves + wave;
    if (q >= nq) return;  // whole wavefront; no workgroup barriers follow
    // Wave-private LDS: keys[lw] | errs[hw] | FP64 values[hw] | rows[hw].
    uint64_t* keys = reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(s_mem) + (size_t)wave * wbytes);
    float* errs = reinterpret_cast<float*>(keys + lw);
    uint64_t* s_v = reinterpret_cast<uint64_t*>(errs + hw);
    uint32_t* s_r = reinterpret_cast<uint32_t*>(s_v + hw);
    uint32_t k, S;
    query_k_s(a, q, k, S);
    uint32_t g = gthr[q];
    uint32_t n = 0;  // keys in the list (wave-uniform)
    const uint64_t cstride = a.cand_off ? a.cand_stride : (uint64_t)nq * cap;
    const uint64_t qbase = a.cand_off ? a.cand_off[q] : (uint64_t)q * cap;

    // Keeps the S smallest keys of keys[0..n) (sorted) and tightens g to the
    // S-th. n <= lw, so the padded length never exceeds lw.
    auto prune = [&]() {
        const uint32_t np = next_pow2(n < 2u ? 2u : n);
        for (uint32_t i = n + lane; i < np; i += kWave) keys[i] = kEmptyKey;
        wave_lds_sync();
        wave_bitonic_sort(keys, np, lane);
        if (n >= S) {
            const uint32_t gs = (uint32_t)(keys[S - 1] >> 32);
            g = gs < g ? gs : g;
            n = S;
        }
        wave_lds_sync();
    };

    // Candidates of up to 64 chunks at a time: lane c holds chunk c's count;
    // entry e belongs to the last chunk whose exclusive prefix is <= e
    // (bisection over the lanes, 6 steps for 64 lanes).
    for (uint32_t c0 = 0; c0 < n_chunks; c0 += kWave) {
        const uint32_t nc = n_chunks - c0 < (uint32_t)kWave ? n_chunks - c0 : (uint32_t)kWave;
        const uint32_t cnt = lane < nc ? counts[(size_t)(c0 + lane) * nq + q] : 0u;
        const uint32_t incl = wave_incl_scan(cnt, lane);
        const uint32_t excl = incl - cnt;
        const uint32_t total = (uint32_t)__shfl((int)incl, kWave - 1);
        for (uint32_t e0 = 0; e0 < total; e0 += 4u * kWave) {
            uint64_t kb[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t e = e0 + (uint32_t)r * kWave + lane;
                const uint32_t ee = e < total ? e : total - 1u;
                uint32_t lo = 0, hi = nc - 1u;
#pragma unroll
                for (int it = 0; it < 6; ++it) {
                    const uint32_t mid = (lo + hi + 1u) >> 1;
                    if ((uint32_t)__shfl((int)excl, (int)mid) <= ee) lo = mid;
                    else hi = mid - 1u;
                }
                const uint32_t base = (uint32_t)__shfl((int)excl, (int)lo);
                kb[r] = e < total ? cand[(c0 + lo) * cstride + qbase + (ee - base)] : kEmptyKey;
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (n + kWave > lw) prune();
                const bool pass = kb[r] != kEmptyKey && (uint32_t)(kb[r] >> 32) <= g;
                const uint64_t bm = __ballot(pass);
                if (pass)
                    keys[n + __builtin_amdgcn_mbcnt_hi((uint32_t)(bm >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u))] = kb[r];
                n += (uint32_t)__popcll(bm);
            }
        }
    }
    wave_lds_sync();
    if (n > S) {
        if (n <= (uint32_t)(kRankSlots * kWave)) {
            // The S smallest (keys are distinct): the S-th smallest key by
            // register bisection, then the keys at or below it, compacted.
            uint64_t mk[kRankSlots];
#pragma unroll
            for (int s2 = 0; s2 < kRankSlots; ++s2) {
                const uint32_t i = (uint32_t)s2 * kWave + lane;
                mk[s2] = i < n ? keys[i] : kEmptyKey;
            }
            const uint64_t t = wave_kth_key<kRankSlots>(mk, S);
            wave_lds_sync();
            uint32_t base = 0;
#pragma unroll
            for (int s2 = 0; s2 < kRankSlots; ++s2) {
                const bool kp = mk[s2] <= t;
                const uint64_t bm = __ballot(kp);
                if (kp)
                    keys[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(bm >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u))] = mk[s2];
                base += (uint32_t)__popcll(bm);
            }
            n = S;
            wave_lds_sync();
        } else {
            prune();
        }
    }
    const uint32_t have = n;
    const uint32_t dim = a.dim;
    const float* qv = Qm + (size_t)q * dim;

    // Bounded rerank: rows that cannot be among the k best are not reranked.
    uint32_t nre = have;  // rows to rerank, first in keys[]
    if (METRIC != METRIC_COSINE && X32 && code_dtype && norms &&
        !(a.flags & (FV_SEARCH_F64_BOUNDARY | FV_SEARCH_FULL_RERANK)) && have > k)
        nre = bounded_rerank_set<METRIC>(keys, have, k, qv, dim, norms, enc_err, code_dtype, lane);

    // Exact rerank in place: keys[i] becomes (distance, row), errs[i] its bound.
    if (X32) {
        const bool vec = (dim & 3u) == 0u && (reinterpret_cast<uintptr_t>(X32) & 15u) == 0u &&
                         (reinterpret_cast<uintptr_t>(Qm) & 15u) == 0u;
        for (uint32_t i0 = 0; i0 < nre; i0 += 2u * kWave) {
            const uint32_t ia = i0 + lane, ib = i0 + kWave + lane;
            // Lanes past the list take a valid row; their results are dropped.
            const uint32_t ra = ia < nre ? key_slot(keys[ia]) : key_slot(keys[0]);
            const uint32_t rb = ib < nre ? key_slot(keys[ib]) : ra;
            float d[2], e[2];
            if (i0 + kWave < nre) {
                const uint32_t rows[2] = {ra, rb};
                const bool live[2] = {ia < nre, ib < nre};
                rerank_rows<METRIC, 2>(X32, qv, dim, vec, rows, live, d, e);
            } else {
                const uint32_t rows[1] = {ra};
                const bool live[1] = {ia < nre};
                float d1[1], e1[1];
                rerank_rows<METRIC, 1>(X32, qv, dim, vec, rows, live, d1, e1);
                d[0] = d1[0];
                e[0] = e1[0];
                d[1] = 0.0f;
                e[1] = 0.0f;
            }
            if (ia < nre) {
                keys[ia] = d[0] < kAdmit ? make_key(d[0], ra) : unadmitted_key(ra);
                errs[ia] = e[0];
            }
            if (ib < nre) {
                keys[ib] = d[1] < kAdmit ? make_key(d[1], rb) : unadmitted_key(rb);
                errs[ib] = e[1];
            }
        }
    } else {
        for (uint32_t i = lane; i < have; i += kWave) {
            const float d = score_as_distance<METRIC>(key_dist(keys[i]), qv, dim);
            keys[i] = d < kAdmit ? make_key(d, key_slot(keys[i])) : unadmitted_key(key_slot(keys[i]));
            errs[i] = 0.0f;
        }
    }
    wave_lds_sync();
    GpuSearchResult* oq = out + (size_t)q * a.top_k;
    const uint32_t kk_out = k < have ? k : have;
    if (!((a.flags & FV_SEARCH_F64_BOUNDARY) && X32 != nullptr && have > 1)) {
        // Without the FP64 boundary only the first k in (distance, row)
        // order are needed: each lane writes its keys of rank < k.
        // Rows left out of the bounded rerank rank after the nre reranked
        // ones (nre >= k), so the first k are among those: the k smallest
        // (bisection to the k-th key), then their ranks.
        uint64_t mk[kRankSlots];
        uint32_t rk[kRankSlots];
        uint32_t nr = nre;
        if (nre > 64u && nre > 2u * kk_out && kk_out > 0 && nre <= (uint32_t)(kRankSlots * kWave)) {
#pragma unroll
            for (int s2 = 0; s2 < kRankSlots; ++s2) {
                const uint32_t i = (uint32_t)s2 * kWave + lane;
                mk[s2] = i < nre ? keys[i] : kEmptyKey;
            }
            const uint64_t t = wave_kth_key<kRankSlots>(mk, kk_out);
            wave_lds_sync();
            uint32_t base = 0;
#pragma unroll
            for (int s2 = 0; s2 < kRankSlots; ++s2) {
                const bool kp = mk[s2] <= t;
                const uint64_t bm = __ballot(kp);
                if (kp)
                    keys[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(bm >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u))] = mk[s2];
                base += (uint32_t)__popcll(bm);
            }
            nr = kk_out;
            wave_lds_sync();
        }
        wave_key_ranks(keys, nr, lane, mk, rk);
#pragma unroll
        for (int s2 = 0; s2 < kRankSlots; ++s2) {
            if (rk[s2] < kk_out) {
                GpuSearchResult r;
                if (key_unadmitted(mk[s2])) {
                    r.id = FV_EMPTY_ID;
                    r.distance = FV_EMPTY_DIST;
                } else {
                    r.id = key_slot(mk[s2]);
                    r.distance = key_dist(mk[s2]);
                }
                oq[rk[s2]] = r;
            }
        }
        for (uint32_t i = kk_out + lane; i < a.top_k; i += kWave) {
            GpuSearchResult r;
            r.id = FV_EMPTY_ID;
            r.distance = FV_EMPTY_DIST;
            oq[i] = r;
        }
        return;
    }
    const uint32_t hp = next_pow2(have < 2u ? 2u : have);
    for (uint32_t i = have + lane; i < hp; i += kWave) keys[i] = kEmptyKey;
    wave_lds_sync();
    wave_bitonic_sort(keys, hp, lane);

    {
        // As in flat_finalize_kernel: candidates adjacent in FP32 order within
        // twice the largest error bound that could reach the top k are
        // re-scored in FP64 (lanes split the dims), then all candidates are
        // ordered by (value, row).
        float em = 0.0f;
        for (uint32_t i = lane; i < have; i += kWave) em = fmaxf(em, errs[i]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) em = fmaxf(em, __shfl_xor(em, off));
        const float band = 2.0000002f * em;
        const uint32_t kk = k < have ? k : have;
        const float lim = key_dist(keys[kk - 1]) + band;
        bool mine = false;
        for (uint32_t i = lane; i < hp; i += kWave) {
            const uint64_t key = keys[i];
            uint32_t r = 0xFFFFFFFFu;
            uint64_t v = ~0ull;
            if (i < have && !key_unadmitted(key)) {
                const float d = key_dist(key);
                r = key_slot(key);
                const bool near_prev = i > 0 && !key_unadmitted(keys[i - 1]) && d - key_dist(keys[i - 1]) <= band;
                const bool near_nxt = i + 1 < have && !key_unadmitted(keys[i + 1]) && key_dist(keys[i + 1]) - d <= band;
                const uint64_t bits = __double_as_longlong((double)d);
                v = (bits & 0x8000000000000000ull) ? ~bits : (bits | 0x8000000000000000ull);
                if ((near_prev || near_nxt) && d <= lim) {
                    r |= 0x80000000u;
                    mine = true;
                }
            }
            s_v[i] = v;
            s_r[i] = r;
        }
        const bool any = __ballot(mine) != 0;
        wave_lds_sync();
        if (any) {
            for (uint32_t i0 = 0; i0 < have; i0 += kWave) {
                const uint32_t i = i0 + lane;
                const uint32_t ri = i < have ? s_r[i] : 0xFFFFFFFFu;
                uint64_t bm = __ballot(ri != 0xFFFFFFFFu && (ri & 0x80000000u));
                while (bm) {
                    const uint32_t j = i0 + (uint32_t)__builtin_ctzll(bm);
                    bm &= bm - 1;
                    const uint32_t r = s_r[j] & 0x7FFFFFFFu;
                    const float* x = X32 + (size_t)r * dim;
                    double acc = 0.0, nqd = 0.0, nxd = 0.0;
                    for (uint32_t d = lane; d < dim; d += kWave) {
                        const double qd = qv[d], xd = x[d];
                        if (METRIC == METRIC_L2) {
                            const double t = qd - xd;
                            acc = __builtin_fma(t, t, acc);
                        } else {
                            acc = __builtin_fma(qd, xd, acc);
                            if (METRIC == METRIC_COSINE) {
                                nqd = __builtin_fma(qd, qd, nqd);
                                nxd = __builtin_fma(xd, xd, nxd);
                            }
                        }
                    }
                    for (int off = 32; off > 0; off >>= 1) {
                        acc += __shfl_xor(acc, off);
                        if (METRIC == METRIC_COSINE) {
                            nqd += __shfl_xor(nqd, off);
                            nxd += __shfl_xor(nxd, off);
                        }
                    }
                    if (lane == 0) {
                        double dd;
                        if (METRIC == METRIC_L2) dd = acc;
                        else if (METRIC == METRIC_INNER_PRODUCT) dd = -acc;
                        else dd = 1.0 - acc / (sqrt(nqd) * sqrt(nxd) + 1e-8);
                        const uint64_t bits = __double_as_longlong(dd);
                        s_v[j] = (bits & 0x8000000000000000ull) ? ~bits : (bits | 0x8000000000000000ull);
                        s_r[j] = r;
                    }
                }
            }
            wave_lds_sync();
            wave_bitonic_sort_pairs(s_v, s_r, hp, (int)lane);
            for (uint32_t i = lane; i < kk; i += kWave) {
                const uint64_t v = s_v[i];
                const uint64_t bits = (v & 0x8000000000000000ull) ? (v & 0x7fffffffffffffffull) : ~v;
                keys[i] = make_key((float)__longlong_as_double(bits), s_r[i]);
            }
            wave_lds_sync();
        }
    }
    for (uint32_t i = lane; i < a.top_k; i += kWave) {
        const uint64_t key = i < k && i < hp ? keys[i] : kEmptyKey;
        GpuSearchResult r;
        if (key_unadmitted(key)) {
            r.id = FV_EMPTY_ID;
            r.distance = FV_EMPTY_DIST;
        } else {
            r.id = key_slot(key);
            r.distance = key_dist(key);
        }
        oq[i] = r;
    }
}

// Negated queries as 16-bit codes (round to nearest even), rows of `pdq`
// elements, zero padded.
template <int DT>
__device__ __forceinline__ void encode_queries_block(uint32_t blk, const float* __restrict__ Qm,
                                                     uint32_t nq, uint32_t dim, uint32_t pdq,
                                                     uint32_t total_rows, uint16_t* __restrict__ out) {
    const size_t i = (size_t)blk * blockDim.x + threadIdx.x;
    if (i >= (size_t)total_rows * pdq) return;
    const uint32_t q = (uint32_t)(i / pdq), d = (uint32_t)(i % pdq);
    const float v = (q < nq && d < dim) ? -Qm[(size_t)q * dim + d] : 0.0f;
    uint16_t h;
    if (DT == FV_DTYPE_F16) {
        const _Float16 f = (_Float16)v;
        h = __builtin_bit_cast(uint16_t, f);
    } else {
        uint32_t u = __float_as_uint(v);
        u += 0x7fffu + ((u >> 16) & 1u);
        h = (uint16_t)(u >> 16);
    }
    out[i] = h;
}

// FP8 e4m3 rows with one FP32 scale per row: s = max|x| / 448 (1 for an all
// zero row), code = e4m3(sign * x / s), zero padded to out_stride bytes. One
// wavefront per row, 4 codes per lane per store. Rows >= n (query block
// padding) are written as zeros with scale 1.
__global__ __launch_bounds__(256) void encode_fp8_rows_kernel(
    const float* __restrict__ x, uint32_t n, uint32_t dim, uint32_t out_stride,
    uint32_t total_rows, float sign, uint8_t* __restrict__ out, float* __restrict__ scales) {
    const uint32_t r = blockIdx.x * 4 + threadIdx.x / kWave;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    if (r >= total_rows) return;
    const float* row = x + (size_t)r * dim;
    float m = 0.0f;
    if (r < n)
        for (uint32_t d = lane; d < dim; d += kWave) m = fmaxf(m, fabsf(row[d]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    const float s = m > 0.0f ? m / 448.0f : 1.0f;
    uint32_t* o = reinterpret_cast<uint32_t*>(out + (size_t)r * out_stride);
    for (uint32_t d0 = 4 * lane; d0 < out_stride; d0 += 4 * kWave) {
        uint32_t w = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint32_t d = d0 + c;
            const uint8_t code = (r < n && d < dim) ? f32_to_e4m3(sign * row[d] / s) : 0u;
            w |= (uint32_t)code << (8 * c);
        }
        o[d0 / 4] = w;
    }
    if (lane == 0 && scales) scales[r] = s;
}

void launch_encode_fp8_rows(const float* x, uint32_t n, uint32_t dim, uint32_t out_stride,
                            uint32_t total_rows, float sign, uint8_t* out, float* scales,
                            hipStream_t stream) {
    encode_fp8_rows_kernel<<<(total_rows + 3) / 4, 256, 0, stream>>>(x, n, dim, out_stride,
                                                                    total_rows, sign, out, scales);
}

// Same selection for large m (m <= 256 * 64): one workgroup per query, 64
// minima per thread in registers, counts reduced across the workgroup.
__global__ __launch_bounds__(256) void bound_select_block_kernel(const float* __restrict__ minima,
                                                                 uint32_t m, uint32_t nq,
                                                                 FlatSearchArgs a,
                                                                 uint32_t* __restrict__ gthr) {
    constexpr int NR = 64;
    __shared__ uint32_t s_part[2][256 / kWave];
    const uint32_t q = blockIdx.x;
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    uint32_t k, S;
    query_k_s(a, q, k, S);
    uint32_t o[NR];
    uint32_t valid = 0;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const uint32_t i = (uint32_t)r * 256u + threadIdx.x;
        uint32_t v = 0xFFFFFFFFu;
        if (i < m) {
            const float f = minima[((size_t)(i >> 4) * nq + q) * 16 + (i & 15u)];
            if (f < kAdmit) v = ord_f32(f);
        }
        o[r] = v;
        valid += v != 0xFFFFFFFFu;
    }
    int par = 0;
    auto block_count = [&](uint32_t c) -> uint32_t {
        for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
        if (lane == 0) s_part[par][wave] = c;
        __syncthreads();
        uint32_t t = 0;
        for (uint32_t w = 0; w < 256 / kWave; ++w) t += s_part[par][w];
        par ^= 1;  // alternate buffers: the next write cannot race this read
        return t;
    };
    if (block_count(valid) < S) return;
    uint32_t lo = 0u, hi = 0xFFFFFFFEu;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        uint32_t c = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) c += o[r] <= mid;
        c = block_count(c);
        if (c >= S) {
            hi = mid;
            if (c <= S + (S >> 3) + 16u) break;
        } else {
            lo = mid + 1u;
        }
    }
    if (threadIdx.x == 0) atomicMin(&gthr[q], hi);
}

// Per-query candidate buffers: query q holds round_up(S_q + slack, 8) keys
// of every chunk at off[q] (exclusive prefix sum by one workgroup, each
// thread summing a contiguous run). Offsets are clamped to the planned keys
// per chunk, so a query whose buffer does not fit (the caller's k sum was
// too small) gets fewer than S_q + slack keys and the scan keeps none of its
// candidates.
__global__ void cand_offsets_kernel(FlatSearchArgs a, uint32_t nq, uint32_t slack, uint32_t budget,
                                    uint32_t* __restrict__ off) {
    __shared__ uint64_t s_part[256];
    auto qkeys = [&](uint32_t q) -> uint64_t {
        uint32_t k, S;
        query_k_s(a, q, k, S);
        return (uint64_t)((S + slack + 7u) & ~7u);
    };
    const uint32_t per = (nq + 255) / 256;
    const uint32_t b0 = threadIdx.x * per;
    uint64_t sum = 0;
    for (uint32_t i = 0; i < per && b0 + i < nq; ++i) sum += qkeys(b0 + i);
    s_part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t run = 0;
        for (int i = 0; i < 256; ++i) {
            const uint64_t x = s_part[i];
            s_part[i] = run;
            run += x;
        }
        off[nq] = (uint32_t)(run < budget ? run : budget);
    }
    __syncthreads();
    uint64_t run = s_part[threadIdx.x];
    for (uint32_t i = 0; i < per && b0 + i < nq; ++i) {
        off[b0 + i] = (uint32_t)(run < budget ? run : budget);
        run += qkeys(b0 + i);
    }
}

// Exclusive prefix sum of n counts by one workgroup (each thread sums a
// contiguous run) into out; the total goes to *total.
__global__ void exclusive_scan_kernel(const uint32_t* __restrict__ v, uint32_t* __restrict__ out,
                                      uint32_t n, uint32_t* __restrict__ total) {
    __shared__ uint32_t s_part[256];
    const uint32_t per = (n + 255) / 256;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t i = 0; i < per && b0 + i < n; ++i) sum += v[b0 + i];
    s_part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < 256; ++i) {
            const uint32_t x = s_part[i];
            s_part[i] = run;
            run += x;
        }
        *total = run;
    }
    __syncthreads();
    uint32_t run = s_part[threadIdx.x];
    for (uint32_t i = 0; i < per && b0 + i < n; ++i) {
        out[b0 + i] = run;
        run += v[b0 + i];
    }
}

// ---- mask compaction: ascending list of the rows whose bit is set ---------

// Valid rows of the mask words of block `blk` (blockDim.x words per block).
__device__ __forceinline__ void mask_block_counts_block(uint32_t blk, const uint32_t* __restrict__ mask,
                                                       uint32_t nwords, uint32_t n,
                                                       uint32_t* __restrict__ block_counts) {
    __shared__ uint32_t s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const uint32_t w = blk * blockDim.x + threadIdx.x;
    uint32_t c = 0;
    if (w < nwords) {
        uint32_t bits = mask[w];
        if ((w + 1) * 32u > n) bits &= (n - w * 32u) >= 32u ? 0xFFFFFFFFu : ((1u << (n - w * 32u)) - 1u);
        c = __popc(bits);
    }
    atomicAdd(&s, c);
    __syncthreads();
    if (threadIdx.x == 0) block_counts[blk] = s;
}

__global__ void mask_block_counts_kernel(const uint32_t* __restrict__ mask, uint32_t nwords,
                                         uint32_t n, uint32_t* __restrict__ block_counts) {
    mask_block_counts_block(blockIdx.x, mask, nwords, n, block_counts);
}

// Search set-up of the 16-bit path in one launch (256-thread blocks): blocks
// [0, nb_q) encode the queries, the next nb_g reset the shared bounds, the
// rest count the valid rows of the mask per block of words. The three jobs
// are independent; one launch instead of three.
template <int DT>
__global__ __launch_bounds__(256) void prep16_kernel(const float* __restrict__ Qm, uint32_t nq,
                                                     uint32_t dim, uint32_t pdq, uint32_t total_rows,
                                                     uint16_t* __restrict__ qh, uint32_t nb_q,
                                                     uint32_t* __restrict__ gthr, uint32_t nb_g,
                                                     const uint32_t* __restrict__ mask, uint32_t nwords,
                                                     uint32_t n, uint32_t* __restrict__ block_counts) {
    const uint32_t b = blockIdx.x;
    if (b < nb_q) {
        encode_queries_block<DT>(b, Qm, nq, dim, pdq, total_rows, qh);
    } else if (b < nb_q + nb_g) {
        const uint32_t i = (b - nb_q) * blockDim.x + threadIdx.x;
        if (i < nq) gthr[i] = 0xFFFFFFFFu;
    } else {
        mask_block_counts_block(b - nb_q - nb_g, mask, nwords, n, block_counts);
    }
}

__global__ void mask_scan_kernel(uint32_t* __restrict__ block_counts, uint32_t nblocks,
                                 uint32_t* __restrict__ total) {
    // One workgroup: exclusive scan of the per-block counts, in place.
    __shared__ uint32_t s_part[256];
    const uint32_t per = (nblocks + 255) / 256;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t i = 0; i < per && b0 + i < nblocks; ++i) sum += block_counts[b0 + i];
    s_part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < 256; ++i) {
            const uint32_t v = s_part[i];
            s_part[i] = run;
            run += v;
        }
        *total = run;
    }
    __syncthreads();
    uint32_t run = s_part[threadIdx.x];
    for (uint32_t i = 0; i < per && b0 + i < nblocks; ++i) {
        const uint32_t v = block_counts[b0 + i];
        block_counts[b0 + i] = run;
        run += v;
    }
}

__global__ void mask_scatter_kernel(const uint32_t* __restrict__ mask, uint32_t nwords,
                                    uint32_t n, const uint32_t* __restrict__ block_offsets,
                                    uint32_t* __restrict__ rows) {
    __shared__ uint32_t s_c[256];
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t bits = 0;
    if (w < nwords) {
        bits = mask[w];
        if ((w + 1) * 32u > n) bits &= (n - w * 32u) >= 32u ? 0xFFFFFFFFu : ((1u << (n - w * 32u)) - 1u);
    }
    s_c[threadIdx.x] = __popc(bits);
    __syncthreads();
    // Inclusive Hillis-Steele scan within the block.
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? s_c[threadIdx.x - off] : 0u;
        __syncthreads();
        s_c[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t pos = block_offsets[blockIdx.x] + s_c[threadIdx.x] - __popc(bits);
    while (bits) {
        const int b = __builtin_ctz(bits);
        bits &= bits - 1;
        rows[pos++] = w * 32u + (uint32_t)b;
    }
}

// ---- planning and launch -----------------------------------------------------

static uint32_t deepest_shortlist(const FlatSearchArgs& a) {
    uint32_t smax = a.factor ? a.factor * a.top_k : (a.shortlist ? a.shortlist : 8u * a.top_k);
    if (smax < a.top_k) smax = a.top_k;
    if (smax > FV_MAX_SHORTLIST) smax = FV_MAX_SHORTLIST;
    return smax;
}

// K steps of one row: 32 elements (16-bit) or 128 (FP8) per step, rounded up
// to the instantiated counts. 0 = unsupported width.
static uint32_t plan_ks(uint32_t padded_dim, int dtype) {
    if (dtype == FV_DTYPE_FP8_E4M3) {
        const uint32_t ks = (padded_dim + 127) / 128;
        if (ks <= 2) return ks;
        for (uint32_t c : {4u, 6u, 8u, 12u, 16u, 24u})
            if (ks <= c) return c;
        return 0;
    }
    const uint32_t ks = (padded_dim + 31) / 32;
    if (ks <= 4) return ks;
    for (uint32_t c : {6u, 8u, 12u, 16u, 24u, 32u, 48u})
        if (ks <= c) return c;
    return 0;
}

FlatMfmaPlan plan_flat_mfma(uint32_t n, uint32_t padded_dim, int dtype, uint32_t nq,
                            const FlatSearchArgs& a, bool has_mask, uint32_t valid_count) {
    FlatMfmaPlan p{};
    const uint32_t smax = deepest_shortlist(a);
    p.s_max = smax;
    // Buffers hold S plus one tile of appends plus one wavefront of bisection
    // slack (tiles have at most 64 rows). A 1024-key buffer absorbs the whole
    // warm-up of a typical chunk, so most (query, chunk) pairs prune at most once.
    if (smax + 2 * 64u <= 1024u) {
        p.cap = 1024u;
        p.prune_regs = 16;
    } else {
        p.cap = next_pow2(smax + 2 * 64u);
        p.prune_regs = 0;  // workgroup prunes through shared scratch
    }
    p.block_prune = p.prune_regs == 0;
    // Deep shortlists with per-query k and a known k sum: each query's
    // buffers hold its own shortlist plus kPerQuerySlack keys (64 of
    // bisection slack and 512 rows between buffer checks), laid out by a
    // prefix sum on the device, instead of the deepest shortlist's
    // power-of-two capacity for every query. The kernels see cap = s_max +
    // slack, which sets their check interval to 512 rows.
    p.percap = false;
    if (p.block_prune && a.d_top_k && a.top_k_sum) {
        // Shortlists are at most F * k + C (k >= 1, so top_k_sum + nq
        // covers k = 0 entries), and never deeper than smax.
        const uint64_t F = a.factor ? a.factor : (a.shortlist ? 1u : 8u);
        const uint64_t C = (!a.factor && a.shortlist) ? a.shortlist : 0u;
        uint64_t ssum = F * ((uint64_t)a.top_k_sum + nq) + C * nq;
        if (ssum > (uint64_t)nq * smax) ssum = (uint64_t)nq * smax;
        const uint64_t keys = ssum + (uint64_t)nq * (kPerQuerySlack + 7u);
        if (keys < (uint64_t)nq * p.cap && keys <= 0xFFFFFFFFull) {
            p.percap = true;
            p.slack = kPerQuerySlack;
            p.cand_keys = keys;
            p.cap = smax + kPerQuerySlack;
        }
    }
    // Per wave: prune scratch and staging list share one region; workgroup
    // prunes use one shared scratch followed by four staging lists.
    p.lds_scan = p.block_prune ? (size_t)kScanSelBytes + 4 * (size_t)kRecBytes
                               : 4 * (size_t)kRecBytes;
    // Shallow shortlists: one fused wavefront per query after the scan.
    p.tail = smax <= kTailMaxS;
    p.tail_lw = next_pow2(smax + kWave);
    p.tail_hw = next_pow2(smax < 2u ? 2u : smax);
    p.tail_wbytes = p.tail_lw * 8 + p.tail_hw * 4;
    if (a.flags & FV_SEARCH_F64_BOUNDARY) p.tail_wbytes += p.tail_hw * (8 + 4);
    p.tail_wbytes = (p.tail_wbytes + 15u) & ~15u;
    p.lds_tail = (size_t)kTailWaves * p.tail_wbytes;
    p.sel_batch = (smax >= 2048u && next_pow2(smax + kSelectBatchLarge) <= 8192u) ? kSelectBatchLarge
                                                                                  : kSelectBatchSmall;
    p.lcap = next_pow2(smax + p.sel_batch);
    p.final_threads = smax >= 2048u ? (uint32_t)kFinalThreads : 256u;
    p.lds_select = (size_t)p.lcap * 8;
    p.hcap = next_pow2(smax < 2u ? 2u : smax);
    p.lds_final = (size_t)p.hcap * 8;
    if (a.flags & FV_SEARCH_F64_BOUNDARY) p.lds_final += (size_t)p.hcap * (8 + 4);
    p.ks = plan_ks(padded_dim, dtype);
    if (p.ks == 0) return p;
    p.qw = (uint32_t)scan_qw(dtype, p.ks);
    p.qb = 4 * p.qw;
    p.qrow_bytes = (uint32_t)scan_row_bytes(dtype, p.ks);
    p.qblocks = (nq + p.qb - 1) / p.qb;
    p.dyn_rows = has_mask && valid_count == FV_VALID_COUNT_DEVICE;
    p.compact = p.dyn_rows || (has_mask && valid_count > 0 && (uint64_t)valid_count * 2 <= n);
    // Device-counted masks plan for every row (the upper bound of the count).
    const uint32_t rows_plan = p.compact && !p.dyn_rows ? valid_count : n;
    // First pass: 16 subset minima per query per chunk over the first
    // bound_rows positions (a quarter of the rows, a sixteenth for rows wider
    // than 256 bytes, at least 64 MB of rows or all). With m >= 4*S minima (at least
    // 256, a power of two, so at least 16 chunks) the S-th smallest
    // sits near rank S of those rows, so the main pass admits about
    // S * rows / bound_rows rows per query from the start.
    uint32_t brows = p.qrow_bytes <= 256 ? rows_plan / 4 : rows_plan / 16;
    // Floor of 64 MB of 256-byte rows, scaled to the row width.
    const uint32_t floor_rows = (uint32_t)((262144ull * 256u) / p.qrow_bytes);
    const uint32_t bmin = rows_plan < floor_rows ? rows_plan : floor_rows;
    if (brows < bmin) brows = bmin;
    if (a.bound_div && a.bound_div != 0xFFFFFFFFu) brows = rows_plan / a.bound_div;
    p.bound_rows = brows;
    uint32_t m = 256u;
    while (m < 4u * smax && m < 16384u) m *= 2u;
    uint32_t bc = m / 16u;
    while (bc > 8u && (uint64_t)bc * 64u > brows) bc /= 2u;  // at least one tile per chunk
    p.bound_chunks = (16u * bc >= smax && brows >= bc * 64u) ? bc : 0u;
    if (a.bound_div == 0xFFFFFFFFu) p.bound_chunks = 0u;
    // Chunks: enough workgroups to cover 256 CUs several times, but at least
    // 8192 rows each without a first pass (every chunk pays a cold start of
    // about S*ln(rows/S) inserts before its threshold settles) or 2048 with
    // one (the seeded bound removes most of it); a multiple of 8 (one slice
    // per XCD) when there are at least 8.
    const uint32_t min_rows = p.bound_chunks ? 2048u : 8192u;
    const uint32_t want = (2560u + p.qblocks - 1) / p.qblocks;
    uint32_t c = chunks_for_rows(rows_plan, want, min_rows, 0xFFFFFFFFu);
    p.chunk_want = p.chunk_min_rows = 0;
    if (a.n_chunks) c = a.n_chunks < 4096u ? a.n_chunks : 4096u;
    else if (p.dyn_rows) {
        p.chunk_want = want;
        p.chunk_min_rows = min_rows;
    }
    p.n_chunks = c;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    p.off_gthr = take((size_t)nq * 4);
    p.off_counts = take((size_t)p.n_chunks * nq * 4);
    if (!p.percap) p.cand_keys = (uint64_t)nq * p.cap;
    p.off_cand = take((size_t)p.n_chunks * p.cand_keys * 8);
    if (p.percap) p.off_candoff = take((size_t)(nq + 1) * 4);
    p.off_qh = take((size_t)p.qblocks * p.qb * p.qrow_bytes);
    p.off_sel = take((size_t)nq * smax * 8);
    p.off_seln = take((size_t)nq * 4);
    p.off_seloff = take((size_t)(nq + 1) * 4);
    if (a.flags & FV_SEARCH_F64_BOUNDARY) p.off_err = take((size_t)nq * smax * 4);
    if (dtype == FV_DTYPE_FP8_E4M3) p.off_qscale = take((size_t)p.qblocks * p.qb * 4);
    if (p.bound_chunks) p.off_minima = take((size_t)p.bound_chunks * nq * 16 * 4);
    if (p.compact) {
        const uint32_t nwords = (n + 31) / 32;
        p.mask_blocks = (nwords + 255) / 256;
        p.off_rows = take((size_t)n * 4);
        p.off_nrows = take(4);
        p.off_blocks = take((size_t)p.mask_blocks * 4);
    }
    p.workspace_bytes = off;
    return p;
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_bound_select(const ScanLaunch& L) {
    const uint32_t m = L.p->bound_chunks * 16;
    const dim3 g((L.nq + 3) / 4);
    if (m <= 1024u) bound_select_kernel<16><<<g, 256, 0, L.stream>>>(L.minima, m, L.nq, L.a, L.gthr);
    else if (m <= 4096u) bound_select_kernel<64><<<g, 256, 0, L.stream>>>(L.minima, m, L.nq, L.a, L.gthr);
    else if (m <= 16384u)
        bound_select_block_kernel<<<L.nq, 256, 0, L.stream>>>(L.minima, m, L.nq, L.a, L.gthr);
    else  // one wavefront per workgroup, minima in LDS (m * 4 bytes, m padded to whole wavefronts)
        bound_select_kernel<0><<<L.nq, kWave, (size_t)((m + kWave - 1) / kWave) * kWave * 4, L.stream>>>(
            L.minima, m, L.nq, L.a, L.gthr);
}

hipError_t launch_flat_mfma(const FvFlatTables* t, const float* Qm, uint32_t nq,
                            const FlatSearchArgs& a, GpuSearchResult* out, void* workspace,
                            const FlatMfmaPlan& p, hipStream_t stream) {
    char* ws = static_cast<char*>(workspace);
    uint32_t* gthr = reinterpret_cast<uint32_t*>(ws + p.off_gthr);
    uint32_t* counts = reinterpret_cast<uint32_t*>(ws + p.off_counts);
    uint64_t* cand = reinterpret_cast<uint64_t*>(ws + p.off_cand);
    const uint32_t* rows = nullptr;
    const uint32_t* nrows = nullptr;
    uint8_t* Qh = reinterpret_cast<uint8_t*>(ws + p.off_qh);
    float* qscale = nullptr;
    const uint32_t qrows = p.qblocks * p.qb;
    uint32_t* blk = reinterpret_cast<uint32_t*>(ws + p.off_blocks);
    const uint32_t nwords = (t->num_vectors + 31) / 32;
    const uint32_t nb_g = (nq + 255) / 256;
    const uint32_t nb_m = p.compact ? p.mask_blocks : 0u;
    if (t->dtype == FV_DTYPE_FP8_E4M3) {
        fill_u32_kernel<<<nb_g, 256, 0, stream>>>(gthr, nq, 0xFFFFFFFFu);
        if (nb_m) mask_block_counts_kernel<<<nb_m, 256, 0, stream>>>(a.mask, nwords, t->num_vectors, blk);
        qscale = reinterpret_cast<float*>(ws + p.off_qscale);
        launch_encode_fp8_rows(Qm, nq, t->dim, p.qrow_bytes, qrows, -1.0f, Qh, qscale, stream);
    } else {
        const uint32_t pdq = p.qrow_bytes / 2;
        const size_t total = (size_t)qrows * pdq;
        const uint32_t nb_q = (uint32_t)((total + 255) / 256);
        uint16_t* q16 = reinterpret_cast<uint16_t*>(Qh);
        const dim3 g(nb_q + nb_g + nb_m);
        if (t->dtype == FV_DTYPE_F16)
            prep16_kernel<FV_DTYPE_F16><<<g, 256, 0, stream>>>(Qm, nq, t->dim, pdq, qrows, q16, nb_q, gthr,
                                                              nb_g, a.mask, nwords, t->num_vectors, blk);
        else
            prep16_kernel<FV_DTYPE_BF16><<<g, 256, 0, stream>>>(Qm, nq, t->dim, pdq, qrows, q16, nb_q, gthr,
                                                               nb_g, a.mask, nwords, t->num_vectors, blk);
    }
    if (p.compact) {
        uint32_t* r = reinterpret_cast<uint32_t*>(ws + p.off_rows);
        uint32_t* nr = reinterpret_cast<uint32_t*>(ws + p.off_nrows);
        mask_scan_kernel<<<1, 256, 0, stream>>>(blk, p.mask_blocks, nr);
        mask_scatter_kernel<<<p.mask_blocks, 256, 0, stream>>>(a.mask, nwords, t->num_vectors, blk, r);
        rows = r;
        nrows = nr;
    }
    float* minima = p.bound_chunks ? reinterpret_cast<float*>(ws + p.off_minima) : nullptr;
    const uint32_t esz = t->dtype == FV_DTYPE_FP8_E4M3 ? 1u : 2u;
    FlatSearchArgs sa = a;
    sa.chunk_want = p.chunk_want;
    sa.chunk_min_rows = p.chunk_min_rows;
    sa.cand_off = nullptr;
    sa.cand_stride = p.cand_keys;
    if (p.percap) {
        uint32_t* co = reinterpret_cast<uint32_t*>(ws + p.off_candoff);
        cand_offsets_kernel<<<1, 256, 0, stream>>>(a, nq, p.slack, (uint32_t)p.cand_keys, co);
        sa.cand_off = co;
    }
    ScanLaunch L{&p, static_cast<const uint8_t*>(t->d_codes), t->padded_dim * esz, t->d_norms_sq,
                 t->d_row_scales, Qh, qscale, nq, sa, rows, nrows, t->num_vectors, gthr, counts,
                 cand, minima, stream};
    hipError_t e = hipErrorInvalidValue;
    const bool large = p.ks > 8, wide = p.ks > 24;
    if (t->dtype == FV_DTYPE_F16)
        e = wide ? scan_f16_wide(L) : (large ? scan_f16_large(L) : scan_f16_small(L));
    else if (t->dtype == FV_DTYPE_BF16)
        e = wide ? scan_bf16_wide(L) : (large ? scan_bf16_large(L) : scan_bf16_small(L));
    else if (t->dtype == FV_DTYPE_FP8_E4M3)
        e = p.ks > 6 ? scan_fp8_wide(L) : scan_fp8(L);
    if (e != hipSuccess) return e;
    const float* X32 = t->d_vectors;
    if (p.tail) {
        const dim3 tg((nq + kTailWaves - 1) / kTailWaves);
        const unsigned tb = kTailWaves * kWave;
        // Bounded rerank for 16-bit codes (scores in distance units).
        const uint32_t cdt = (t->dtype == FV_DTYPE_F16 || t->dtype == FV_DTYPE_BF16) ? t->dtype : 0u;
        if (a.metric == METRIC_L2)
            flat_tail_kernel<METRIC_L2><<<tg, tb, p.lds_tail, stream>>>(
                cand, counts, p.n_chunks, nq, p.cap, gthr, sa, p.tail_lw, p.tail_hw, p.tail_wbytes, X32, Qm, out,
                t->d_norms_sq, t->d_enc_err, cdt);
        else if (a.metric == METRIC_INNER_PRODUCT)
            flat_tail_kernel<METRIC_INNER_PRODUCT><<<tg, tb, p.lds_tail, stream>>>(
                cand, counts, p.n_chunks, nq, p.cap, gthr, sa, p.tail_lw, p.tail_hw, p.tail_wbytes, X32, Qm, out,
                t->d_norms_sq, t->d_enc_err, cdt);
        else
            flat_tail_kernel<METRIC_COSINE><<<tg, tb, p.lds_tail, stream>>>(
                cand, counts, p.n_chunks, nq, p.cap, gthr, sa, p.tail_lw, p.tail_hw, p.tail_wbytes, X32, Qm, out,
                t->d_norms_sq, t->d_enc_err, 0u);
        return hipGetLastError();
    }
    uint64_t* sel = reinterpret_cast<uint64_t*>(ws + p.off_sel);
    uint32_t* sel_n = reinterpret_cast<uint32_t*>(ws + p.off_seln);
    uint32_t* sel_off = reinterpret_cast<uint32_t*>(ws + p.off_seloff);
    float* err = (a.flags & FV_SEARCH_F64_BOUNDARY) ? reinterpret_cast<float*>(ws + p.off_err) : nullptr;
    // Bounded rerank set for 16-bit codes with FP32 rows (L2 and inner
    // product), unless every selected row is to be reranked. With the FP64
    // boundary the rows left out are beyond k kept rows in FP64 order too
    // (the intervals bound the exact distance), so ids and order are those of
    // the full rerank; the near-tie band is twice the largest error bound of
    // the reranked rows, so which kept rows report an FP64-rounded distance
    // can differ from a full rerank (as it does between shortlist depths).
    SelectBound sb{Qm, t->d_norms_sq, t->d_enc_err, t->dim, 0u, a.metric};
    if ((t->dtype == FV_DTYPE_F16 || t->dtype == FV_DTYPE_BF16) && X32 && a.metric != METRIC_COSINE &&
        !(a.flags & FV_SEARCH_FULL_RERANK))
        sb.code_dtype = t->dtype;
    if (p.sel_batch == kSelectBatchLarge)
        flat_select_kernel<kSelectBatchLarge><<<nq, kSelectBatchLarge, p.lds_select, stream>>>(
            cand, counts, p.n_chunks, nq, p.cap, gthr, sa, p.lcap, sel, sel_n, sb);
    else
        flat_select_kernel<kSelectBatchSmall><<<nq, kSelectBatchSmall, p.lds_select, stream>>>(
            cand, counts, p.n_chunks, nq, p.cap, gthr, sa, p.lcap, sel, sel_n, sb);
    // Compact work list for the rerank: exclusive prefix of the per-query
    // counts (the last slot holds the total).
    exclusive_scan_kernel<<<1, 256, 0, stream>>>(sel_n, sel_off, nq, sel_off + nq);
    const dim3 rg((unsigned)(((size_t)nq * p.s_max + kRerankRows - 1) / kRerankRows));
    // Finalize. With per-query k and deep shortlists, queries of at most
    // kFinalSplitS rows go to a launch of 256-thread workgroups with LDS for
    // that many keys (many per CU) and only the deeper ones to the
    // 1024-thread launch; each launch skips the other's queries.
    const bool split = a.d_top_k != nullptr && p.final_threads == (uint32_t)kFinalThreads;
    const uint32_t hs = kFinalSplitS;
    const size_t lds_small = (size_t)hs * 8 + ((a.flags & FV_SEARCH_F64_BOUNDARY) ? (size_t)hs * 12 : 0);
    auto finalize = [&](auto kern) {
        if (split) {
            kern<<<nq, 256, lds_small, stream>>>(sel, sel_n, err, hs, a, X32, Qm, out, 0u, hs + 1u);
            kern<<<nq, p.final_threads, p.lds_final, stream>>>(sel, sel_n, err, p.hcap, a, X32, Qm, out,
                                                               hs + 1u, 0xFFFFFFFFu);
        } else {
            kern<<<nq, p.final_threads, p.lds_final, stream>>>(sel, sel_n, err, p.hcap, a, X32, Qm, out,
                                                               0u, 0xFFFFFFFFu);
        }
    };
    if (a.metric == METRIC_L2) {
        flat_rerank_kernel<METRIC_L2><<<rg, 256, 0, stream>>>(X32, Qm, t->dim, nq, p.s_max, sel_n, sel_off, sel_off + nq, sel, err);
        finalize(flat_finalize_kernel<METRIC_L2>);
    } else if (a.metric == METRIC_INNER_PRODUCT) {
        flat_rerank_kernel<METRIC_INNER_PRODUCT><<<rg, 256, 0, stream>>>(X32, Qm, t->dim, nq, p.s_max, sel_n, sel_off, sel_off + nq, sel, err);
        finalize(flat_finalize_kernel<METRIC_INNER_PRODUCT>);
    } else {
        flat_rerank_kernel<METRIC_COSINE><<<rg, 256, 0, stream>>>(X32, Qm, t->dim, nq, p.s_max, sel_n, sel_off, sel_off + nq, sel, err);
        finalize(flat_finalize_kernel<METRIC_COSINE>);
    }
    return hipGetLastError();
}

// Widest rows the matrix-core scan takes: 1536 dims of 16-bit codes, 3072
// of FP8 (16 queries per wave at the widest).
static uint32_t max_scan_dim(uint32_t dtype) {
    return dtype == FV_DTYPE_FP8_E4M3 ? 3072u : 1536u;
}

static int check_tables(const FvFlatTables* t) {
    if (!t->d_codes || !t->d_norms_sq) return -1;
    if (t->dtype == FV_DTYPE_FP8_E4M3 && !t->d_row_scales) return -1;
    if (t->dtype != FV_DTYPE_F16 && t->dtype != FV_DTYPE_BF16 && t->dtype != FV_DTYPE_FP8_E4M3)
        return (int)hipErrorInvalidValue;
    if (t->padded_dim == 0 || t->padded_dim % 16 || t->dim == 0 || t->dim > t->padded_dim ||
        t->padded_dim > max_scan_dim(t->dtype))
        return (int)hipErrorInvalidValue;
    return 0;
}

static bool make_args(const FvFlatTables* t, const FvFlatSearchParams* sp, FlatSearchArgs& a) {
    a = FlatSearchArgs{};
    a.mask = sp->d_valid_mask;
    a.d_top_k = sp->d_top_k;
    a.top_k = sp->top_k;
    a.factor = sp->shortlist_factor;
    a.shortlist = sp->shortlist;
    a.metric = (int)sp->metric;
    a.dim = t->dim;
    a.flags = sp->flags;
    a.n_chunks = sp->n_chunks;
    a.bound_div = sp->bound_div;
    a.top_k_sum = sp->d_top_k ? sp->top_k_sum : 0u;
    if (a.top_k == 0 || a.top_k > FV_MAX_TOPK) return false;
    if (sp->metric > METRIC_INNER_PRODUCT ||
        (sp->flags & ~(FV_SEARCH_F64_BOUNDARY | FV_SEARCH_FULL_RERANK)) != 0)
        return false;
    a.s_max = deepest_shortlist(a);
    return true;
}

bool flat_plan_for(const FvFlatTables* t, uint32_t nq, const FvFlatSearchParams* sp,
                   FlatMfmaPlan* out) {
    if (!t || !sp || !out || t->num_vectors == 0 || check_tables(t) != 0) return false;
    FlatSearchArgs a;
    if (!make_args(t, sp, a)) return false;
    *out = plan_flat_mfma(t->num_vectors, t->padded_dim, (int)t->dtype, nq, a,
                          sp->d_valid_mask != nullptr, sp->valid_count);
    return out->ks != 0 && out->lcap <= 8192u;
}

}  // namespace fv

extern "C" {

size_t fv_flat_workspace_size(const FvFlatTables* tables, uint32_t num_queries,
                              const FvFlatSearchParams* params) {
    if (!tables || !params || tables->num_vectors == 0) return 0;
    fv::FlatSearchArgs a;
    if (!fv::make_args(tables, params, a)) return 0;
    if (fv::check_tables(tables) != 0) return 0;
    fv::FlatMfmaPlan p = fv::plan_flat_mfma(tables->num_vectors, tables->padded_dim,
                                            (int)tables->dtype, num_queries, a,
                                            params->d_valid_mask != nullptr, params->valid_count);
    if (p.lcap > 8192u || p.ks == 0) return 0;
    return p.workspace_bytes;
}

int fv_flat_search(const FvFlatTables* tables, const float* d_queries, uint32_t num_queries,
                   const FvFlatSearchParams* params, GpuSearchResult* d_out_results,
                   void* d_workspace, size_t workspace_bytes, void* stream) {
    if (!tables || !d_queries || !d_out_results || !params) return -1;
    const int tc = fv::check_tables(tables);
    if (tc) return tc;
    fv::FlatSearchArgs a;
    if (!fv::make_args(tables, params, a)) return (int)hipErrorInvalidValue;
    if (num_queries == 0) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (tables->num_vectors == 0)
        return (int)fv::fill_results(d_out_results, (size_t)num_queries * a.top_k, FV_EMPTY_ID,
                                     0x7149f2cau /* 1e30f */, s);
    fv::FlatMfmaPlan p = fv::plan_flat_mfma(tables->num_vectors, tables->padded_dim,
                                            (int)tables->dtype, num_queries, a,
                                            params->d_valid_mask != nullptr, params->valid_count);
    if (p.lcap > 8192u || p.ks == 0) return (int)hipErrorInvalidValue;  // shortlist deeper than 7936
    if ((a.flags & FV_SEARCH_F64_BOUNDARY) && tables->num_vectors > 0x7FFFFFFFu)
        return (int)hipErrorInvalidValue;  // the boundary pass marks rows with bit 31
    if (!d_workspace || workspace_bytes < p.workspace_bytes) return (int)hipErrorInvalidValue;
    hipError_t e = fv::launch_flat_mfma(tables, d_queries, num_queries, a, d_out_results,
                                        d_workspace, p, s);
    return e == hipSuccess ? 0 : (int)e;
}

// cuda_bridge.h:138-147 / kernels/src/ivf_quant_lookup.cu:1106-1123.
int cuda_fp16_tc_search_batch(const Fp16GpuTables* tables, const float* d_queries,
                              uint32_t num_queries, uint32_t top_k, MetricType metric,
                              const uint32_t* d_valid_mask, GpuSearchResult* d_out_results,
                              void* stream) {
    if (!tables || !d_queries || !d_out_results) return -1;
    if (num_queries == 0 || top_k == 0 || tables->num_vectors == 0 || tables->padded_dim == 0)
        return 0;  // ivf_quant_lookup.cu:1047-1049: silent no-op
    if (top_k > 256u) top_k = 256u;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (tables->padded_dim > fv::max_scan_dim(FV_DTYPE_F16) && tables->d_vectors &&
        tables->dim > 0 && tables->dim <= tables->padded_dim) {
        // Rows wider than the matrix-core scan takes: the exact FP32 scan of
        // the FP32 rows, which returns what the exact rerank of a complete
        // shortlist returns (the k best by exact distance, ties to the lower row).
        const fv::FlatExactPlan ep = fv::plan_flat_exact(tables->num_vectors, tables->dim,
                                                         num_queries, top_k, (int)metric);
        fv::WorkspaceLease ws;
        hipError_t e = fv::stream_workspace(s, ep.workspace_bytes, ws);
        if (e != hipSuccess) return (int)e;
        e = fv::launch_flat_exact(tables->d_vectors, nullptr, tables->num_vectors, tables->dim,
                                  d_queries, num_queries, top_k, (int)metric, d_valid_mask,
                                  d_out_results, top_k, ws.ptr, ep, s, false);
        return e == hipSuccess ? 0 : (int)e;
    }
    FvFlatTables t{};
    t.d_codes = tables->d_vectors_fp16;
    t.d_norms_sq = tables->d_vec_norms_sq;
    t.d_vectors = tables->d_vectors;
    t.num_vectors = tables->num_vectors;
    t.dim = tables->dim;
    t.padded_dim = tables->padded_dim;
    t.dtype = FV_DTYPE_F16;
    const int tc = fv::check_tables(&t);
    if (tc) return tc;
    FvFlatSearchParams sp{};
    sp.d_valid_mask = d_valid_mask;
    // The mask's popcount is taken on the device: sparse masks (the 5%-valid
    // filters of the benchmarks) are compacted and only their rows scored.
    sp.valid_count = d_valid_mask ? FV_VALID_COUNT_DEVICE : 0u;
    sp.metric = (uint32_t)metric;
    sp.top_k = top_k;
    fv::FlatSearchArgs a;
    if (!fv::make_args(&t, &sp, a)) return (int)hipErrorInvalidValue;
    fv::FlatMfmaPlan p = fv::plan_flat_mfma(t.num_vectors, t.padded_dim, FV_DTYPE_F16, num_queries,
                                            a, d_valid_mask != nullptr, sp.valid_count);
    if (p.lcap > 8192u || p.ks == 0) return (int)hipErrorInvalidValue;
    fv::WorkspaceLease ws;
    hipError_t e = fv::stream_workspace(s, p.workspace_bytes, ws);
    if (e != hipSuccess) return (int)e;
    e = fv::launch_flat_mfma(&t, d_queries, num_queries, a, d_out_results, ws.ptr, p, s);
    return e == hipSuccess ? 0 : (int)e;
}

}  // extern "C"
