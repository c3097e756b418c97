// Backward pass of top-k inner-product retrieval (fv_topk_ip_backward).
//
// No counterpart in the reference, which has no training path. For query i
// and its selected rows r_ij = ids[i][j] (j < k), the scores are
// s_ij = q_i . x_{r_ij}. Given dL/ds (or, in fused mode, the softmax
// cross-entropy of s/tau against a target position t_i, whose gradient is
// scale * (softmax(s_i/tau)_j - [j == t_i]) / tau), the kernel forms
//   dL/dq_i    = sum_j g_ij x_{r_ij}            (gather of the k rows)
//   dL/dx_r   += sum_{(i,j): r_ij = r} g_ij q_i  (scatter-add, float atomics)
// and never materialises a Q x N score matrix: the fused mode recomputes the
// k scores of each query from Q, the rows and the ids.
//
// One workgroup per query: the query row in LDS; in fused mode each wave
// computes scores of rows j = wave, wave+4, ... with lanes splitting the dims
// (fixed reduction tree, so the scores are deterministic); softmax in LDS;
// dQ by one thread per dim (rows read coalesced); dX by atomic adds, one
// thread per dim per selected row. Empty ids (0xFFFFFFFF) are skipped.
#include "fv_flat_exact.hpp"

namespace fv {

constexpr uint32_t kBwdMaxDim = 4096;
constexpr uint32_t kBwdMaxK = 1024;

__global__ __launch_bounds__(256) void topk_ip_backward_kernel(
    const float* __restrict__ Q, const float* __restrict__ X, uint32_t n, uint32_t dim,
    const uint32_t* __restrict__ ids, uint32_t k, const float* __restrict__ dS,
    const uint32_t* __restrict__ targets, float inv_tau, float scale,
    const float* __restrict__ grad_out, float* __restrict__ dQ, float* __restrict__ dX,
    float* __restrict__ scores_out) {
    __shared__ float s_q[kBwdMaxDim];
    __shared__ float s_g[kBwdMaxK];
    __shared__ uint32_t s_r[kBwdMaxK];
    __shared__ float s_red[2];
    const uint32_t i = blockIdx.x;
    const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const float* qi = Q + (size_t)i * dim;
    for (uint32_t d = tid; d < dim; d += blockDim.x) s_q[d] = qi[d];
    for (uint32_t j = tid; j < k; j += blockDim.x) {
        const uint32_t r = ids[(size_t)i * k + j];
        s_r[j] = r < n ? r : 0xFFFFFFFFu;
    }
    __syncthreads();
    if (dS) {
        for (uint32_t j = tid; j < k; j += blockDim.x)
            s_g[j] = s_r[j] != 0xFFFFFFFFu ? dS[(size_t)i * k + j] : 0.0f;
    } else {
        // Recompute the scores: one wave per selected row.
        for (uint32_t j = wave; j < k; j += blockDim.x / kWave) {
            const uint32_t r = s_r[j];
            float p = 0.0f;
            if (r != 0xFFFFFFFFu) {
                const float* x = X + (size_t)r * dim;
                for (uint32_t d = lane; d < dim; d += kWave) p = __builtin_fmaf(s_q[d], x[d], p);
            }
            for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off);
            if (lane == 0) s_g[j] = r != 0xFFFFFFFFu ? p : -__builtin_inff();
        }
        __syncthreads();
        if (scores_out)
            for (uint32_t j = tid; j < k; j += blockDim.x) scores_out[(size_t)i * k + j] = s_g[j];
        // Softmax of s / tau over the valid entries (wave 0), then the
        // cross-entropy gradient against the target position.
        if (wave == 0) {
            float m = -__builtin_inff();
            for (uint32_t j = lane; j < k; j += kWave) m = fmaxf(m, s_g[j] * inv_tau);
            for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
            float z = 0.0f;
            for (uint32_t j = lane; j < k; j += kWave)
                z += s_r[j] != 0xFFFFFFFFu ? expf(s_g[j] * inv_tau - m) : 0.0f;
            for (int off = 32; off > 0; off >>= 1) z += __shfl_xor(z, off);
            if (lane == 0) {
                s_red[0] = m;
                s_red[1] = z;
            }
        }
        __syncthreads();
        const float m = s_red[0], z = s_red[1];
        const uint32_t t = targets ? targets[i] : 0xFFFFFFFFu;
        if (grad_out) scale *= *grad_out;
        for (uint32_t j = tid; j < k; j += blockDim.x) {
            float g = 0.0f;
            if (s_r[j] != 0xFFFFFFFFu) {
                const float p = expf(s_g[j] * inv_tau - m) / z;
                g = scale * (p - (j == t ? 1.0f : 0.0f)) * inv_tau;
            }
            s_g[j] = g;
        }
    }
    __syncthreads();
    // dQ_i = sum_j g_ij x_{r_ij}, sequential over j.
    for (uint32_t d = tid; d < dim; d += blockDim.x) {
        float acc = 0.0f;
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t r = s_r[j];
            if (r != 0xFFFFFFFFu) acc = __builtin_fmaf(s_g[j], X[(size_t)r * dim + d], acc);
        }
        dQ[(size_t)i * dim + d] = acc;
    }
    // dX_r += g_ij q_i.
    if (dX) {
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t r = s_r[j];
            const float g = s_g[j];
            if (r == 0xFFFFFFFFu || g == 0.0f) continue;
            float* xr = dX + (size_t)r * dim;
            for (uint32_t d = tid; d < dim; d += blockDim.x) unsafeAtomicAdd(xr + d, g * s_q[d]);
        }
    }
}

}  // namespace fv

extern "C" int fv_topk_ip_backward(const float* d_queries, uint32_t num_queries,
                                   const float* d_vectors, uint32_t num_vectors, uint32_t dim,
                                   const uint32_t* d_ids, uint32_t top_k,
                                   const float* d_grad_scores, const uint32_t* d_targets,
                                   float inv_tau, float loss_scale, const float* d_grad_output,
                                   float* d_grad_queries, float* d_grad_vectors,
                                   float* d_scores_out, void* stream) {
    if (!d_queries || !d_vectors || !d_ids || !d_grad_queries) return -1;
    if (!d_grad_scores && !d_targets) return -1;
    if (dim == 0 || dim > fv::kBwdMaxDim || top_k == 0 || top_k > fv::kBwdMaxK)
        return (int)hipErrorInvalidValue;
    if (num_queries == 0) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    fv::topk_ip_backward_kernel<<<num_queries, 256, 0, s>>>(
        d_queries, d_vectors, num_vectors, dim, d_ids, top_k, d_grad_scores, d_targets, inv_tau,
        loss_scale, d_grad_output, d_grad_queries, d_grad_vectors, d_scores_out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}
