// Device encoders for the flat tables (the step before the scan).
//
// Reference: crates/engine/src/quant.rs:306-327 (encode_fp16_padded: rows
// padded to a multiple of 16 halves, |x|^2 of the FP32 originals). Here the
// encoder runs on the GPU, rounds to nearest even (the reference rounds half
// up on magnitude; the two differ only on exact binary16 ties) and also emits
// BF16 codes, FP8 e4m3 codes with per-row scales, and FP64 norms.
#include "fv_flat_mfma.hpp"

namespace fv {

__global__ void encode_16bit_kernel(const float* __restrict__ x, uint32_t n, uint32_t dim,
                                    uint32_t pdim, uint32_t dtype, uint16_t* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * pdim) return;
    uint32_t r = (uint32_t)(i / pdim), d = (uint32_t)(i % pdim);
    float v = d < dim ? x[(size_t)r * dim + d] : 0.0f;
    uint16_t h;
    if (dtype == FV_DTYPE_F16) {
        h = f32_to_f16_rows(v);
    } else {
        uint32_t u = __float_as_uint(v);
        u += 0x7fffu + ((u >> 16) & 1u);
        h = (uint16_t)(u >> 16);
    }
    out[i] = h;
}

__global__ void norms_kernel(const float* __restrict__ x, uint32_t n, uint32_t dim,
                             float* __restrict__ out32, double* __restrict__ out64) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float* row = x + (size_t)r * dim;
    float a = 0.0f;
    double b = 0.0;
    auto add = [&](float v) {
        float sq = v * v;
        a = a + sq;
        b = __builtin_fma((double)v, (double)v, b);
    };
    if ((dim & 3u) == 0u && (reinterpret_cast<uintptr_t>(row) & 15u) == 0u) {
        // Same sequential order, 128-bit loads.
        const float4* r4 = reinterpret_cast<const float4*>(row);
        for (uint32_t i = 0; i < dim / 4; ++i) {
            const float4 v = r4[i];
            add(v.x);
            add(v.y);
            add(v.z);
            add(v.w);
        }
    } else {
        for (uint32_t d = 0; d < dim; ++d) add(row[d]);
    }
    if (out32) out32[r] = a;
    if (out64) out64[r] = b;
}

// |decode(code_row) - x_row| per row: FP64 sum of squared differences, its
// square root rounded up to FP32 (a non-finite difference gives +inf).
__global__ void encode_error_kernel(const float* __restrict__ x, const uint16_t* __restrict__ codes,
                                    uint32_t n, uint32_t dim, uint32_t pdim, uint32_t dtype,
                                    float* __restrict__ out) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float* row = x + (size_t)r * dim;
    const uint16_t* c = codes + (size_t)r * pdim;
    double acc = 0.0;
    for (uint32_t d = 0; d < dim; ++d) {
        const float dec = dtype == FV_DTYPE_F16 ? (float)__builtin_bit_cast(_Float16, c[d])
                                                : __uint_as_float((uint32_t)c[d] << 16);
        const double diff = (double)dec - (double)row[d];
        acc = __builtin_fma(diff, diff, acc);
    }
    float e = (float)sqrt(acc);
    if (!(e < __builtin_inff())) e = __builtin_inff();
    else if ((double)e < sqrt(acc)) e = __uint_as_float(__float_as_uint(e) + 1u);
    out[r] = e;
}

}  // namespace fv

extern "C" int fv_encode_error(const float* d_vectors, const void* d_codes, uint32_t num_vectors,
                               uint32_t dim, uint32_t padded_dim, uint32_t dtype, float* d_err_out,
                               void* stream) {
    if (!d_vectors || !d_codes || !d_err_out) return -1;
    if (dtype != FV_DTYPE_F16 && dtype != FV_DTYPE_BF16) return (int)hipErrorInvalidValue;
    if (padded_dim < dim || dim == 0) return (int)hipErrorInvalidValue;
    if (num_vectors == 0) return 0;
    fv::encode_error_kernel<<<(num_vectors + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(
        d_vectors, static_cast<const uint16_t*>(d_codes), num_vectors, dim, padded_dim, dtype,
        d_err_out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

extern "C" int fv_encode_flat(const float* d_vectors, uint32_t num_vectors, uint32_t dim,
                              uint32_t padded_dim, uint32_t dtype, void* d_codes_out,
                              float* d_norms_out, double* d_norms_f64_out,
                              float* d_row_scales_out, void* stream) {
    if (!d_vectors) return -1;
    if (padded_dim < dim || dim == 0) return (int)hipErrorInvalidValue;
    if (num_vectors == 0) return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (d_codes_out && dtype == FV_DTYPE_FP8_E4M3) {
        if (padded_dim % 16) return (int)hipErrorInvalidValue;
        fv::launch_encode_fp8_rows(d_vectors, num_vectors, dim, padded_dim, num_vectors, 1.0f,
                                   static_cast<uint8_t*>(d_codes_out), d_row_scales_out, s);
    } else if (d_codes_out) {
        if (dtype != FV_DTYPE_F16 && dtype != FV_DTYPE_BF16) return (int)hipErrorInvalidValue;
        size_t total = (size_t)num_vectors * padded_dim;
        fv::encode_16bit_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(
            d_vectors, num_vectors, dim, padded_dim, dtype, static_cast<uint16_t*>(d_codes_out));
    }
    if (d_norms_out || d_norms_f64_out)
        fv::norms_kernel<<<(num_vectors + 255) / 256, 256, 0, s>>>(d_vectors, num_vectors, dim,
                                                                  d_norms_out, d_norms_f64_out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}
