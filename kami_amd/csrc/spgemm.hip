// Block-sparse (BSR) SpGEMM C = A * B: symbolic (count, scan, fill) and numeric phases.
//
// Reference: src/sparse/spgemm/2d/block_spgemm_2d_half_mma.cu -- spgemm_symbolic_kernel
// (:58-86, one thread per block row with a dense flag array), thrust::exclusive_scan
// (:806), a blocking copy of nnzC (:808), spgemm_symbolic_col_kernel (:88-116, ascending
// column fill) and the numeric kernel block_spgemm_2d_half_mma (:171) that keeps the whole
// matrix in one CTA.  Here every phase is its own device kernel sized for large matrices:
//   count  -- one workgroup per block row; the row's output columns are a bitmap in LDS
//             set with LDS atomics while walking B's rows named by A's columns; the
//             count is the bitmap's population;
//   scan   -- one workgroup scans the counts in place into rowptrC (exclusive, with
//             rowptrC[mb] = nnzb(C)); chunks of 1024 with a running carry;
//   fill   -- the bitmap is rebuilt and its set bits are written in ascending order,
//             each thread placing the bits of one 32-bit word after the prefix popcount
//             of the words before it;
//   numeric -- one wave per output block (i, j): the row i comes from a 64-way search
//             of rowptrC; the lanes binary-search B's sorted rows k of all blocks (i, k)
//             of A at once for column j; A fragments are loaded straight into
//             registers, B blocks (row-major K x N) are transposed through the wave's LDS
//             into the K-contiguous MFMA layout.
// All integer phases are exact and deterministic.  B may have at most 491520 block
// columns (a 60 KiB bitmap).
#include <algorithm>

#include "kami_internal.hpp"

namespace kami {

constexpr int kMaxBitmapWords = 15360;  // 491520 block columns: a 60 KiB bitmap

// Mark the output columns of block row `row` in the LDS bitmap (bitmap must be zero).
__device__ inline void mark_row(int row, const int32_t* rpA, const int32_t* ciA,
                                const int32_t* rpB, const int32_t* ciB, unsigned* bitmap) {
    const int a0 = rpA[row], a1 = rpA[row + 1];
    for (int t = a0; t < a1; ++t) {
        const int k = ciA[t];
        const int b0 = rpB[k], b1 = rpB[k + 1];
        for (int l = b0 + (int)threadIdx.x; l < b1; l += blockDim.x) {
            const int j = ciB[l];
            atomicOr(&bitmap[j >> 5], 1u << (j & 31));
        }
    }
}

__global__ void __launch_bounds__(256)
spgemm_count_kernel(int mb, int nwords, const int32_t* rpA, const int32_t* ciA,
                    const int32_t* rpB, const int32_t* ciB, int32_t* counts) {
    extern __shared__ unsigned bitmap[];
    __shared__ int s_part[8];
    for (int row = blockIdx.x; row < mb; row += gridDim.x) {
        for (int w = threadIdx.x; w < nwords; w += blockDim.x) bitmap[w] = 0u;
        __syncthreads();
        mark_row(row, rpA, ciA, rpB, ciB, bitmap);
        __syncthreads();
        int c = 0;
        for (int w = threadIdx.x; w < nwords; w += blockDim.x) c += __popc(bitmap[w]);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
        if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int q = 0; q < (int)(blockDim.x >> 6); ++q) tot += s_part[q];
            counts[row] = tot;
        }
        __syncthreads();
    }
}

// In-place exclusive scan of x[0..n-2] into x[0..n-1] (x[n-1] receives the total;
// its input value is ignored).  One workgroup of 1024 threads.
// Exclusive scan of the n - 1 per-row counts in x into x[0 .. n-1].  The running total is
// kept in 64 bits and every prefix is stored saturated at INT32_MAX, so an output with
// 2^31 - 1 or more blocks (which int32 row pointers cannot hold) shows as x[n-1] ==
// INT32_MAX instead of wrapping.
__global__ void __launch_bounds__(1024) spgemm_scan_kernel(int32_t* x, int n) {
    __shared__ int s_wave[16];
    __shared__ long long s_carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + tid;
        const int v = (i < n - 1) ? x[i] : 0;
        int incl = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        if (lane == 63) s_wave[wave] = incl;
        __syncthreads();
        if (wave == 0) {
            int w = lane < 16 ? s_wave[lane] : 0;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                const int y = __shfl_up(w, off);
                if (lane >= off) w += y;
            }
            if (lane < 16) s_wave[lane] = w;  // inclusive wave totals
        }
        __syncthreads();
        const long long carry = s_carry;
        const int before = (wave > 0 ? s_wave[wave - 1] : 0) + incl - v;  // < 2^30 per chunk
        if (i < n) x[i] = (int32_t)min(carry + before, (long long)INT32_MAX);
        __syncthreads();
        if (tid == 0) s_carry = carry + s_wave[15];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256)
spgemm_fill_kernel(int mb, int nwords, const int32_t* rpA, const int32_t* ciA,
                   const int32_t* rpB, const int32_t* ciB, const int32_t* rpC, int32_t* ciC) {
    extern __shared__ unsigned bitmap[];
    __shared__ int s_wave[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int row = blockIdx.x; row < mb; row += gridDim.x) {
        for (int w = tid; w < nwords; w += blockDim.x) bitmap[w] = 0u;
        __syncthreads();
        mark_row(row, rpA, ciA, rpB, ciB, bitmap);
        __syncthreads();
        int out = rpC[row];
        for (int base = 0; base < nwords; base += 256) {
            const int w = base + tid;
            const unsigned bits = w < nwords ? bitmap[w] : 0u;
            const int v = __popc(bits);
            int incl = v;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int y = __shfl_up(incl, off);
                if (lane >= off) incl += y;
            }
            if (lane == 63) s_wave[wave] = incl;
            __syncthreads();
            int before = incl - v;
            for (int q = 0; q < wave; ++q) before += s_wave[q];
            int pos = out + before;
            unsigned b = bits;
            while (b) {
                const int bit = __ffs(b) - 1;
                ciC[pos++] = (w << 5) + bit;
                b &= b - 1;
            }
            const int tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
            __syncthreads();
            out += tot;
        }
        __syncthreads();
    }
}

// ---- numeric: one wave per output block ----

template <int BS, class AT>
struct SpNumGeom {
    static constexpr int IM = BS / 16;                 // atom rows (= atom cols)
    static constexpr int KSTEPS = BS / AT::KSTEP;
    static constexpr int ELT = (int)sizeof(typename AT::in_t);
    static constexpr int ROWB = lds_row_bytes(BS * ELT, AT::FRAG_BYTES);  // transposed B row
    static constexpr int LDS = BS * ROWB;
    static constexpr int CHUNKS = BS * BS * ELT / 16;   // 16-byte pieces of a block
    static constexpr int PER_LANE = (CHUNKS + 63) / 64;
};

__device__ inline int find_col(const int32_t* ci, int lo, int hi, int j) {
    // index in [lo, hi) with ci[idx] == j, or -1 (ci ascending within the row)
    const int end = hi;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ci[mid] < j) lo = mid + 1; else hi = mid;
    }
    return (lo < end && ci[lo] == j) ? lo : -1;
}

template <int BS, class AT>
__global__ void __launch_bounds__(64)
spgemm_numeric_kernel(int mb, int64_t nnzc, const int32_t* rpA, const int32_t* ciA,
                      const void* vA_, const int32_t* rpB, const int32_t* ciB, const void* vB_,
                      const int32_t* rpC, const int32_t* ciC, void* vC_, double alpha_d,
                      int iters) {
    using G = SpNumGeom<BS, AT>;
    using in_t = typename AT::in_t;
    using acc_t = typename AT::acc_t;
    using out_t = typename AT::out_t;
    using frag_t = typename AT::frag_t;
    using accv_t = typename AT::accv_t;
    constexpr int FB = AT::FRAG_BYTES;
    __shared__ __attribute__((aligned(16))) unsigned char sB[G::LDS];
    const in_t* vA = reinterpret_cast<const in_t*>(vA_);
    const in_t* vB = reinterpret_cast<const in_t*>(vB_);
    out_t* vC = reinterpret_cast<out_t*>(vC_);
    const int lane = threadIdx.x;
    const acc_t alpha = (acc_t)alpha_d;

    for (int64_t cb = blockIdx.x; cb < nnzc; cb += gridDim.x) {
        // block row of output block cb: the last row with rpC[row] <= cb, found by a
        // 64-way search (each round the lanes probe 64 evenly spaced row pointers)
        int lo = 0, hi = mb;
        while (hi - lo > 1) {
            const int step = (hi - lo + 63) / 64;
            const int idx = lo + (lane + 1) * step;
            const bool le = idx < hi && (int64_t)rpC[idx] <= cb;
            const int c = __popcll(__ballot(le));  // probes are monotone: a prefix is set
            const int nlo = lo + c * step;
            hi = min(hi, nlo + step);
            lo = nlo;
        }
        const int i = lo, j = ciC[cb];
        accv_t acc[G::IM][G::IM];
#pragma unroll
        for (int a = 0; a < G::IM; ++a)
#pragma unroll
            for (int b = 0; b < G::IM; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[a][b][v] = 0;

        // every lane searches the B row of one block of A's row i for column j; the
        // matches are then multiplied in ascending order of A's blocks
        const int a0 = rpA[i], a1 = rpA[i + 1];
        for (int base = a0; base < a1; base += 64) {
          const int tl = base + lane;
          int lmatch = -1;
          if (tl < a1) {
              const int k = ciA[tl];
              lmatch = find_col(ciB, rpB[k], rpB[k + 1], j);
          }
          uint64_t todo = __ballot(lmatch >= 0);
          while (todo) {
            const int src = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            const int t = base + src;
            const int l = __shfl(lmatch, src);
            const in_t* Ablk = vA + (int64_t)t * BS * BS;
            const in_t* Bblk = vB + (int64_t)l * BS * BS;
            // B block (row-major K x N) -> LDS as N rows of K (transposed)
#pragma unroll
            for (int q = 0; q < G::PER_LANE; ++q) {
                const int c = lane + 64 * q;
                if (c >= G::CHUNKS) break;
                constexpr int CPR = BS * G::ELT / 16;   // pieces per block row
                const int kr = c / CPR, jc = (c % CPR) * (16 / G::ELT);
                const u4_t raw = load_raw<16>(Bblk + kr * BS + jc);
                const in_t* e = reinterpret_cast<const in_t*>(&raw);
#pragma unroll
                for (int x = 0; x < 16 / G::ELT; ++x)
                    *reinterpret_cast<in_t*>(sB + (jc + x) * G::ROWB + kr * G::ELT) = e[x];
            }
            __syncthreads();
            // both blocks' fragments in registers; `iters` > 1 repeats the products on them
            // (the reference's in-kernel NUM_ITER repeat), iters = 1 is the plain product
            frag_t af[G::KSTEPS][G::IM], bf[G::KSTEPS][G::IM];
#pragma unroll
            for (int ks = 0; ks < G::KSTEPS; ++ks) {
                const int kk = ks * AT::KSTEP + AT::KV * (lane >> 4);
#pragma unroll
                for (int x = 0; x < G::IM; ++x) {
                    af[ks][x] = bitcast<frag_t>(load_raw<FB>(Ablk + (x * 16 + (lane & 15)) * BS + kk));
                    bf[ks][x] = bitcast<frag_t>(
                        load_raw<FB>(sB + (x * 16 + (lane & 15)) * G::ROWB + kk * G::ELT));
                }
            }
            __syncthreads();
            for (int it = 0; it < iters; ++it)
#pragma unroll
                for (int ks = 0; ks < G::KSTEPS; ++ks)
#pragma unroll
                    for (int jn = 0; jn < G::IM; ++jn)
#pragma unroll
                        for (int im = 0; im < G::IM; ++im) AT::mma(acc[im][jn], af[ks][im], bf[ks][jn]);
          }
        }
        out_t* Cblk = vC + cb * BS * BS;
#pragma unroll
        for (int im = 0; im < G::IM; ++im)
#pragma unroll
            for (int jn = 0; jn < G::IM; ++jn)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    Cblk[(im * 16 + AT::out_row(lane, v)) * BS + jn * 16 + (lane & 15)] =
                        (out_t)(alpha * acc[im][jn][v]);
    }
}

static kami_status check_bsr(const kami_bsr* X) {
    if (!X || X->mb <= 0 || X->nb <= 0 || X->nnzb < 0) return KAMI_EINVAL;
    if (!X->rowptr || (X->nnzb > 0 && (!X->colidx))) return KAMI_EINVAL;
    return KAMI_OK;
}

static kami_status check_pair(const kami_bsr* A, const kami_bsr* B) {
    kami_status st = check_bsr(A);
    if (st == KAMI_OK) st = check_bsr(B);
    if (st != KAMI_OK) return st;
    if (A->nb != B->mb || A->bs != B->bs) return KAMI_EINVAL;
    if (B->nb > kMaxBitmapWords * 32) return KAMI_ENOTSUP;
    return KAMI_OK;
}

static int sym_grid(int mb) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return std::max(1, std::min(mb, cus * 8));
}

}  // namespace kami

using namespace kami;

extern "C" {

kami_status kami_bsr_spgemm_workspace_size(int32_t mb, int32_t nb, size_t* bytes) {
    if (mb <= 0 || nb <= 0 || !bytes) return KAMI_EINVAL;
    *bytes = 0;  // counts are scanned in place in rowptrC
    return KAMI_OK;
}

kami_status kami_bsr_spgemm_symbolic(const kami_bsr* A, const kami_bsr* B, int32_t* rowptrC,
                                     int64_t* nnzbC, void* ws, size_t ws_bytes, void* stream) {
    (void)ws;
    (void)ws_bytes;
    kami_status st = check_pair(A, B);
    if (st != KAMI_OK) return st;
    if (!rowptrC || !nnzbC) return KAMI_EINVAL;
    st = device_check();
    if (st != KAMI_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const int nwords = (B->nb + 31) / 32;
    hipLaunchKernelGGL(spgemm_count_kernel, dim3(sym_grid(A->mb)), dim3(256), nwords * 4, s, A->mb,
                       nwords, A->rowptr, A->colidx, B->rowptr, B->colidx, rowptrC);
    hipLaunchKernelGGL(spgemm_scan_kernel, dim3(1), dim3(1024), 0, s, rowptrC, A->mb + 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e);
    int32_t total = 0;
    e = hipMemcpyAsync(&total, rowptrC + A->mb, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e);
    if (total == INT32_MAX) return KAMI_ENOTSUP;  // more blocks than int32 row pointers hold
    *nnzbC = total;
    return KAMI_OK;
}

kami_status kami_bsr_spgemm_fill(const kami_bsr* A, const kami_bsr* B, const int32_t* rowptrC,
                                 int32_t* colidxC, void* ws, size_t ws_bytes, void* stream) {
    (void)ws;
    (void)ws_bytes;
    kami_status st = check_pair(A, B);
    if (st != KAMI_OK) return st;
    if (!rowptrC || !colidxC) return KAMI_EINVAL;
    st = device_check();
    if (st != KAMI_OK) return st;
    const int nwords = (B->nb + 31) / 32;
    hipLaunchKernelGGL(spgemm_fill_kernel, dim3(sym_grid(A->mb)), dim3(256), nwords * 4,
                       (hipStream_t)stream, A->mb, nwords, A->rowptr, A->colidx, B->rowptr,
                       B->colidx, rowptrC, colidxC);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e);
    return KAMI_OK;
}

kami_status kami_bsr_spgemm_numeric(const kami_bsr* A, const kami_bsr* B, const int32_t* rowptrC,
                                    const int32_t* colidxC, int64_t nnzbC, void* valC,
                                    int32_t precision, double alpha, void* stream) {
    return kami_bsr_spgemm_numeric_repeat(A, B, rowptrC, colidxC, nnzbC, valC, precision, alpha,
                                          1, stream);
}

kami_status kami_bsr_spgemm_numeric_repeat(const kami_bsr* A, const kami_bsr* B,
                                           const int32_t* rowptrC, const int32_t* colidxC,
                                           int64_t nnzbC, void* valC, int32_t precision,
                                           double alpha, int32_t iters, void* stream) {
    if (iters < 1) return KAMI_EINVAL;
    kami_status st = check_pair(A, B);
    if (st != KAMI_OK) return st;
    if (nnzbC < 0 || !rowptrC) return KAMI_EINVAL;
    if (nnzbC == 0) return KAMI_OK;
    if (!colidxC || !valC || !A->vals || !B->vals) return KAMI_EINVAL;
    if (((uintptr_t)A->vals % 16) || ((uintptr_t)B->vals % 16) || ((uintptr_t)valC % 16))
        return KAMI_EINVAL;
    if (A->bs != 16 && A->bs != 32) return KAMI_ENOTSUP;
    st = device_check();
    if (st != KAMI_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const unsigned grid = (unsigned)std::min<int64_t>(nnzbC, 1 << 20);
#define KAMI_SPNUM(BS, AT)                                                                      \
    hipLaunchKernelGGL((spgemm_numeric_kernel<BS, AT>), dim3(grid), dim3(64), 0, s, A->mb,       \
                       nnzbC, A->rowptr, A->colidx, A->vals, B->rowptr, B->colidx, B->vals,      \
                       rowptrC, colidxC, valC, alpha, (int)iters)
    if (precision == KAMI_PREC_F16_F32) {
        if (A->bs == 16) KAMI_SPNUM(16, AtomF16x16); else KAMI_SPNUM(32, AtomF16x32);
    } else if (precision == KAMI_PREC_BF16_F32) {
        if (A->bs == 16) KAMI_SPNUM(16, AtomBF16x16); else KAMI_SPNUM(32, AtomBF16x32);
    } else if (precision == KAMI_PREC_F64) {
        if (A->bs == 16) KAMI_SPNUM(16, AtomF64); else KAMI_SPNUM(32, AtomF64);
    } else {
        return KAMI_ENOTSUP;
    }
#undef KAMI_SPNUM
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e);
    return KAMI_OK;
}

}  // extern "C"
