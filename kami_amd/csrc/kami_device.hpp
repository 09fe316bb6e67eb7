// Device-side building blocks for gfx950 (CDNA4, wave64).
//
// An "atom" is one 16x16 output tile of an MFMA instruction together with the
// fragment a lane holds for one K step.  Lane layouts (verified on MI355X with an
// in-bounds probe):
//   * operand fragments: lane l holds row (l % 16) of A (or column (l % 16) of B) and a
//     run of KV consecutive k values starting at KV * (l / 16) inside the K step.  A and B
//     use the same k assignment, so the product is independent of how the hardware
//     orders k inside the instruction;
//   * FP64 16x16x4 accumulator: element v of lane l is C[l/16 + 4v][l % 16];
//   * FP32 16x16 accumulators:  element v of lane l is C[4(l/16) + v][l % 16].
// FP64 takes two 16x16x4 instructions per 16-byte fragment (KV = 2, K step 8); FP16 and
// BF16 take one 16x16x32 instruction per 16-byte fragment (KV = 8, K step 32) or one
// 16x16x16 instruction per 8-byte fragment (KV = 4, K step 16) when K is only a
// multiple of 16.
//
// These replace the reference's inline-PTX mma.sync wrappers
// (src/block_gemm/KAMI/1d/fp64/block_gemm_1d_double_mma.cu:42-53 and
//  src/block_gemm/KAMI/1d/fp16/block_gemm_1d_half_mma.cu:43-59); the FP16 atoms
// accumulate in FP32 where the reference accumulates in FP16.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <stdint.h>

namespace kami {

constexpr int kWave = 64;

// Byte stride of LDS rows that hold `data` bytes each and are read as MFMA fragments:
// lane l reads `frag` bytes at row (l % 16), offset (l / 16) * frag.  gfx950 serves 32
// lanes of such a read per pass from a 512-byte bank window, so the read is free of bank
// conflicts when the stride is frag * s with s twice an odd number (measured for 16-byte
// fragments: 96-, 160-, 224- and 288-byte rows read at 218-230 B/clk per CU, 144-, 272-
// and 320-byte rows at 120-126; for 8-byte fragments 80- and 144-byte rows are fastest).
constexpr int lds_row_bytes(int data, int frag) {
    int s = (data + frag - 1) / frag;
    while (s % 4 != 2) ++s;
    return s * frag;
}

typedef double d2_t __attribute__((ext_vector_type(2)));
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef _Float16 h4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef short s4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u2_t __attribute__((ext_vector_type(2)));

struct AtomF64 {
    using in_t = double;
    using acc_t = double;
    using out_t = double;
    using frag_t = d2_t;
    using accv_t = d4_t;
    using raw_t = u4_t;
    static constexpr int KV = 2;
    static constexpr int KSTEP = 8;
    static constexpr int FRAG_BYTES = 16;
    static constexpr int NSUB = 2;  // instructions per fragment
    __device__ static inline void mma_sub(accv_t& c, const frag_t& a, const frag_t& b, int e) {
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[e], b[e], c, 0, 0, 0);
    }
    __device__ static inline void mma(accv_t& c, const frag_t& a, const frag_t& b) {
        mma_sub(c, a, b, 0);
        mma_sub(c, a, b, 1);
    }
    __device__ static inline int out_row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

struct AtomF16x32 {
    using in_t = uint16_t;
    using acc_t = float;
    using out_t = float;
    using frag_t = h8_t;
    using accv_t = f4_t;
    using raw_t = u4_t;
    static constexpr int KV = 8;
    static constexpr int KSTEP = 32;
    static constexpr int FRAG_BYTES = 16;
    static constexpr int NSUB = 1;
    __device__ static inline void mma_sub(accv_t& c, const frag_t& a, const frag_t& b, int) {
        mma(c, a, b);
    }
    __device__ static inline void mma(accv_t& c, const frag_t& a, const frag_t& b) {
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
    __device__ static inline int out_row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

struct AtomF16x16 {
    using in_t = uint16_t;
    using acc_t = float;
    using out_t = float;
    using frag_t = h4_t;
    using accv_t = f4_t;
    using raw_t = u2_t;
    static constexpr int KV = 4;
    static constexpr int KSTEP = 16;
    static constexpr int FRAG_BYTES = 8;
    static constexpr int NSUB = 1;
    __device__ static inline void mma_sub(accv_t& c, const frag_t& a, const frag_t& b, int) {
        mma(c, a, b);
    }
    __device__ static inline void mma(accv_t& c, const frag_t& a, const frag_t& b) {
        c = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
    }
    __device__ static inline int out_row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

struct AtomBF16x32 {
    using in_t = uint16_t;
    using acc_t = float;
    using out_t = float;
    using frag_t = bf8_t;
    using accv_t = f4_t;
    using raw_t = u4_t;
    static constexpr int KV = 8;
    static constexpr int KSTEP = 32;
    static constexpr int FRAG_BYTES = 16;
    static constexpr int NSUB = 1;
    __device__ static inline void mma_sub(accv_t& c, const frag_t& a, const frag_t& b, int) {
        mma(c, a, b);
    }
    __device__ static inline void mma(accv_t& c, const frag_t& a, const frag_t& b) {
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static inline int out_row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

struct AtomBF16x16 {
    using in_t = uint16_t;
    using acc_t = float;
    using out_t = float;
    using frag_t = s4_t;
    using accv_t = f4_t;
    using raw_t = u2_t;
    static constexpr int KV = 4;
    static constexpr int KSTEP = 16;
    static constexpr int FRAG_BYTES = 8;
    static constexpr int NSUB = 1;
    __device__ static inline void mma_sub(accv_t& c, const frag_t& a, const frag_t& b, int) {
        mma(c, a, b);
    }
    __device__ static inline void mma(accv_t& c, const frag_t& a, const frag_t& b) {
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    __device__ static inline int out_row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

typedef float f16x_t __attribute__((ext_vector_type(16)));

// 32 x 32 x 16 atoms (gfx950 MFMA, full rate where 16x16x16 runs at half): a wave computes
// a 32 x 32 tile; lane l holds column (or A row) l % 32, k values 8 (l / 32) .. + 7, and
// 16 accumulators, rows 8 (v / 4) + 4 (l / 32) + v % 4 for v < 16.  Only the block-GEMM
// template takes them (it reads DIM and NACC); 16 x 16 atoms leave both out.
struct AtomF16M32 {
    using in_t = uint16_t;
    using acc_t = float;
    using out_t = float;
    using frag_t = h8_t;
    using accv_t = f16x_t;
    using raw_t = u4_t;
    static constexpr int KV = 8;
    static constexpr int KSTEP = 16;
    static constexpr int FRAG_BYTES = 16;
    static constexpr int NSUB = 1;
    static constexpr int DIM = 32, NACC = 16;
    __device__ static inline void mma_sub(accv_t& c, const frag_t& a, const frag_t& b, int) {
        mma(c, a, b);
    }
    __device__ static inline void mma(accv_t& c, const frag_t& a, const frag_t& b) {
        c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
    __device__ static inline int out_row(int lane, int v) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }
};

struct AtomBF16M32 {
    using in_t = uint16_t;
    using acc_t = float;
    using out_t = float;
    using frag_t = bf8_t;
    using accv_t = f16x_t;
    using raw_t = u4_t;
    static constexpr int KV = 8;
    static constexpr int KSTEP = 16;
    static constexpr int FRAG_BYTES = 16;
    static constexpr int NSUB = 1;
    static constexpr int DIM = 32, NACC = 16;
    __device__ static inline void mma_sub(accv_t& c, const frag_t& a, const frag_t& b, int) {
        mma(c, a, b);
    }
    __device__ static inline void mma(accv_t& c, const frag_t& a, const frag_t& b) {
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static inline int out_row(int lane, int v) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }
};

// Edge of an atom's output tile and accumulators per lane: 16 and 4 unless the atom
// names them (the 32 x 32 atoms).
template <class AT, class = void>
struct AtomShape {
    static constexpr int DIM = 16, NACC = 4;
};
template <class AT>
struct AtomShape<AT, std::void_t<decltype(AT::DIM)>> {
    static constexpr int DIM = AT::DIM, NACC = AT::NACC;
};

// K step 8 on the 16x16x16 instruction, for the 3D split of K = 16 over two depths: lane
// groups 0 and 1 hold the depth's eight k values; groups 2 and 3 (whose fragment loads
// fall outside the depth's slice) contribute zeros.
template <class Base>
struct AtomK8 : Base {
    static constexpr int KSTEP = 8;
    __device__ static inline void mma_sub(typename Base::accv_t& c, const typename Base::frag_t& a,
                                          const typename Base::frag_t& b, int) {
        mma(c, a, b);
    }
    __device__ static inline void mma(typename Base::accv_t& c, const typename Base::frag_t& a,
                                      const typename Base::frag_t& b) {
        const bool live = (__lane_id() >> 5) == 0;
        const typename Base::frag_t z = {};
        Base::mma(c, live ? a : z, live ? b : z);
    }
};
using AtomF16x16K8 = AtomK8<AtomF16x16>;
using AtomBF16x16K8 = AtomK8<AtomBF16x16>;

template <class To, class From>
__device__ inline To bitcast(const From& f) {
    static_assert(sizeof(To) == sizeof(From), "bitcast size");
    return __builtin_bit_cast(To, f);
}

// 16- or 8-byte loads/stores used for fragments and staging.
template <int BYTES>
struct Raw;
template <>
struct Raw<16> {
    using type = u4_t;
};
template <>
struct Raw<8> {
    using type = u2_t;
};

template <int BYTES>
__device__ inline typename Raw<BYTES>::type load_raw(const void* p) {
    return *reinterpret_cast<const typename Raw<BYTES>::type*>(p);
}
template <int BYTES>
__device__ inline void store_raw(void* p, const typename Raw<BYTES>::type& v) {
    *reinterpret_cast<typename Raw<BYTES>::type*>(p) = v;
}

}  // namespace kami
