// Entry points beside the GEMM paths: the digit slicing of the split-precision path as a
// standalone kernel, and the reference programs' known-answer inputs generated in device
// memory.
//
// kami_split_slices writes the same digits the split GEMM kernels form in LDS
// (split_f64.hip): per row and per 32-wide chunk of K an exponent e with every |x| of the
// chunk below 2^e (0 for an all-zero chunk), then r = x 2^(BETA - e) and, slice by slice,
// d = rint(r) (round to nearest, ties to even), r = (r - d) 2^BETA.
//
// kami_generate fills a device array with the inputs of the reference's programs:
//   ramp: x[i] = i  (src/block_gemm/KAMI/1d/fp64/block_gemm_1d_double_mma.cu:224-236);
//   rand_r(&seed) % 3 with seed = i, per element
//         (src/batched_gemm/KAMI/1d/batched_gemm_1d_double_mma.cu:231-245).
// rand_r is the C library's reentrant generator; glibc defines it as three steps of the
// linear congruential recurrence next = next * 1103515245 + 12345 (mod 2^32), taking 11,
// 10 and 10 bits of next / 65536 into a 31-bit result.  The device restates that
// definition; the tests compare it with the host library's rand_r.
#include <algorithm>
#include <cstdint>

#include "kami_internal.hpp"

namespace kami {

// ---- slicing ----

struct SliceArgs {
    const double* X;
    uint16_t* digits;
    int32_t* exps;
    int64_t rows, k, ldx, plane;  // plane = rows * k (elements of one digit plane)
    int32_t nchunk, slices;
};

__device__ __forceinline__ int slice_exp(double x) {
    return x == 0.0 ? -4096 : __builtin_amdgcn_frexp_exp(x);
}

// Workgroups of 256 threads cover 256 columns of a row (rows strided by gridDim.y);
// lane l of wave w handles column 256 b + 64 w + l.  The 32 lanes of a half-wave form
// one chunk: its maximum exponent is reduced with xor shuffles inside the half-wave.
template <int BETA, bool BF16>
__global__ void __launch_bounds__(256) split_slices_kernel(SliceArgs a) {
    const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool in = col < a.k;
    for (int64_t row = blockIdx.y; row < a.rows; row += gridDim.y) {  // uniform
        const double x = in ? a.X[row * a.ldx + col] : 0.0;
        int e = in ? slice_exp(x) : -4096;
#pragma unroll
        for (int off = 16; off >= 1; off >>= 1) e = max(e, __shfl_xor(e, off, 64));
        if (e == -4096) e = 0;
        if (!in) continue;
        if ((col & 31) == 0) a.exps[row * a.nchunk + (col >> 5)] = e;
        double r = __builtin_amdgcn_ldexp(x, BETA - e);
        uint16_t* out = a.digits + row * a.k + col;
        for (int s = 0; s < a.slices; ++s) {
            const double q = __builtin_rint(r);
            // |q| <= 2^BETA: exact in float, in binary16 and in bfloat16
            const float f = (float)q;
            uint16_t bits;
            if constexpr (BF16)
                bits = (uint16_t)(__builtin_bit_cast(uint32_t, f) >> 16);
            else
                bits = __builtin_bit_cast(uint16_t, (_Float16)f);
            out[s * a.plane] = bits;
            r = (r - q) * (double)(1 << BETA);
        }
    }
}

// ---- known-answer generators ----

__device__ __forceinline__ uint32_t rand_r_step(uint32_t& next) {
    next = next * 1103515245u + 12345u;
    return next / 65536u;
}

__device__ __forceinline__ int rand_r_value(uint32_t seed) {
    uint32_t next = seed;
    uint32_t result = rand_r_step(next) % 2048u;
    result = (result << 10) ^ (rand_r_step(next) % 1024u);
    result = (result << 10) ^ (rand_r_step(next) % 1024u);
    return (int)result;
}

struct GenArgs {
    void* dst;
    int64_t count, offset;
    int32_t kind, precision;
};

__global__ void __launch_bounds__(256) generate_kernel(GenArgs a) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.count; i += stride) {
        const int64_t j = i + a.offset;
        const double v = a.kind == KAMI_GEN_RAMP ? (double)j
                                                 : (double)(rand_r_value((uint32_t)j) % 3);
        if (a.precision == KAMI_PREC_F16_F32) {
            // through float: exact below 2^24, then round to nearest even
            static_cast<uint16_t*>(a.dst)[i] = __builtin_bit_cast(uint16_t, (_Float16)(float)v);
        } else if (a.precision == KAMI_PREC_BF16_F32) {
            // round to nearest even on the upper 16 bits of the float
            const uint32_t u = __builtin_bit_cast(uint32_t, (float)v);
            static_cast<uint16_t*>(a.dst)[i] = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
        } else {
            static_cast<double*>(a.dst)[i] = v;
        }
    }
}

}  // namespace kami

using namespace kami;

extern "C" {

kami_status kami_split_slices(const double* X, int64_t rows, int64_t k, int64_t ldx,
                              int32_t precision, int32_t slices, uint16_t* digits,
                              int32_t* exps, void* stream) {
    if (!X || !digits || !exps || rows < 0 || k < 0) return KAMI_EINVAL;
    if (precision != KAMI_PREC_F64_SPLIT_F16 && precision != KAMI_PREC_F64_SPLIT_BF16)
        return KAMI_ENOTSUP;
    if (slices < 1 || slices > 8) return KAMI_EINVAL;
    if (ldx == 0) ldx = k;
    if (ldx < k) return KAMI_EINVAL;
    if (rows == 0 || k == 0) return KAMI_OK;
    if (!span_fits(rows, ldx, 8) || !span_fits(rows, k, 2 * (int64_t)slices)) return KAMI_EINVAL;
    // columns span one grid dimension of 256-thread workgroups, below 2^32 work-items
    if ((k + 255) / 256 * 256 >= ((int64_t)1 << 32)) return KAMI_ENOTSUP;
    kami_status st = device_check();
    if (st != KAMI_OK) return st;
    SliceArgs a{X, digits, exps, rows, k, ldx, rows * k, (int32_t)((k + 31) / 32), slices};
    const dim3 grid((unsigned)((k + 255) / 256), (unsigned)std::min<int64_t>(rows, 65535));
    hipStream_t s = (hipStream_t)stream;
    if (precision == KAMI_PREC_F64_SPLIT_F16)
        hipLaunchKernelGGL((split_slices_kernel<9, false>), grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((split_slices_kernel<8, true>), grid, dim3(256), 0, s, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? KAMI_OK : hip_fail(e);
}

kami_status kami_generate(int32_t kind, int32_t precision, void* dst, int64_t count,
                          int64_t offset, void* stream) {
    if (!dst || count < 0 || offset < 0) return KAMI_EINVAL;
    if (kind != KAMI_GEN_RAMP && kind != KAMI_GEN_RAND_R3) return KAMI_EINVAL;
    if (precision != KAMI_PREC_F64 && precision != KAMI_PREC_F16_F32 &&
        precision != KAMI_PREC_BF16_F32)
        return KAMI_ENOTSUP;
    const size_t align = precision == KAMI_PREC_F64 ? 8 : 2;
    if ((uintptr_t)dst % align) return KAMI_EINVAL;
    if (count == 0) return KAMI_OK;
    // the values' indices offset + i stay inside int64
    if (!span_fits(count, 1, (int64_t)align) || offset > INT64_MAX - count) return KAMI_EINVAL;
    kami_status st = device_check();
    if (st != KAMI_OK) return st;
    GenArgs a{dst, count, offset, kind, precision};
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 65536);
    hipLaunchKernelGGL(generate_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? KAMI_OK : hip_fail(e);
}

}  // extern "C"
