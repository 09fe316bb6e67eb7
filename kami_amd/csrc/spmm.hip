// Block SpMM: C (M x N) = alpha * A * op(B) + beta * C with A in BSR and B dense.
//
// The reference computes one such product per CTA with A re-tiled per warp rank
// (src/sparse/spmm/2d/block_spmm_2d_half_mma.cu:72-235, bsr2tilebsr :264-291) and B stored
// N x K, so C(i, j) = sum_k A(i, k) B[j * K + k] (check loop :378-403).  Here A stays in
// plain BSR (the same format the SpGEMM entry points take) and the product is spread
// over workgroups:
//   * one workgroup per (block row i, 64-column tile of C); wave w owns the 16 columns
//     [16 w, 16 w + 16) of the tile and all bs rows of the block row;
//   * for every stored block (i, kb) the workgroup stages the A block (bs x bs) and the
//     B slice (64 columns x bs values of K, from B stored N x K or K x N) in LDS, rows
//     padded to the conflict-free stride; the next block is loaded into registers while the current one
//     is multiplied (double-buffered LDS, one barrier per block);
//   * FP16 / BF16 inputs accumulate in FP32, FP64 in FP64; block rows without stored
//     blocks still write beta * C.
#include <algorithm>

#include "kami_internal.hpp"

namespace kami {

template <int BS, class AT>
struct SpmmGeom {
    using in_t = typename AT::in_t;
    static constexpr int NT = 256, TN = 64;  // C tile: BS rows x TN columns, 4 waves
    static constexpr int ELT = (int)sizeof(in_t), VE = 16 / ELT;
    static constexpr int ROWB = lds_row_bytes(BS * ELT, AT::FRAG_BYTES);
    static constexpr int A_BYTES = BS * ROWB;
    static constexpr int STAGE = (BS + TN) * ROWB;
    static constexpr int CPR = BS / VE;      // 16-byte pieces per staged row
    static constexpr int NA = BS * CPR;      // pieces of the A block
    static constexpr int NB = TN * CPR;      // pieces of the B slice (either layout)
    static constexpr int NP = (NA + NB + NT - 1) / NT;
    static constexpr int IM = BS / 16;
    static_assert(BS % AT::KSTEP == 0 && BS % 16 == 0, "block size");
    static_assert(TN == 16 * (NT / 64), "one 16-column atom per wave");
};

struct SpmmArgs {
    int32_t mb, nb, n, b_kn;
    int64_t nnzb, ldb, ldc;
    const int32_t* rowptr;
    const int32_t* colidx;
    const void* vals;
    const void* B;
    void* C;
    double alpha, beta;
};

template <int BS, class AT>
__global__ void __launch_bounds__(256) spmm_kernel(SpmmArgs a) {
    using G = SpmmGeom<BS, AT>;
    using in_t = typename AT::in_t;
    using acc_t = typename AT::acc_t;
    using out_t = typename AT::out_t;
    using frag_t = typename AT::frag_t;
    using accv_t = typename AT::accv_t;
    constexpr int FB = AT::FRAG_BYTES;
    __shared__ __attribute__((aligned(16))) unsigned char smem[2 * G::STAGE];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = blockIdx.x;                     // block row of A and C
    const int n0 = blockIdx.y * G::TN;            // first column of the C tile
    const bool wave_live = n0 + wave * 16 < a.n;  // wave-uniform
    const in_t* __restrict__ vals = reinterpret_cast<const in_t*>(a.vals);
    const in_t* __restrict__ B = reinterpret_cast<const in_t*>(a.B);
    // stored blocks of row i (clamped to the arrays the caller described)
    const int64_t rb = a.rowptr[i], re = a.rowptr[i + 1];
    const int64_t beg = rb < 0 ? 0 : rb;
    const int64_t end = re > a.nnzb ? a.nnzb : re;

    u4_t r[G::NP];
    auto load = [&](int64_t q) {
        const int kb = a.colidx[q];
        const bool kb_ok = kb >= 0 && kb < a.nb;
#pragma unroll
        for (int p = 0; p < G::NP; ++p) {
            const int idx = tid + p * G::NT;
            u4_t v = {0, 0, 0, 0};
            if (idx < G::NA) {
                v = load_raw<16>(vals + q * BS * BS + (idx / G::CPR) * BS + (idx % G::CPR) * G::VE);
            } else if (idx < G::NA + G::NB && kb_ok) {
                const int j = idx - G::NA;
                if (!a.b_kn) {
                    const int nn = j / G::CPR, c = j % G::CPR;
                    if (n0 + nn < a.n)
                        v = load_raw<16>(B + (int64_t)(n0 + nn) * a.ldb + (int64_t)kb * BS + c * G::VE);
                } else {
                    const int kr = j / (G::TN / G::VE), jc = (j % (G::TN / G::VE)) * G::VE;
                    if (n0 + jc < a.n)
                        v = load_raw<16>(B + ((int64_t)kb * BS + kr) * a.ldb + n0 + jc);
                }
            }
            r[p] = v;
        }
    };
    auto write = [&](int buf) {
        unsigned char* base = smem + buf * G::STAGE;
#pragma unroll
        for (int p = 0; p < G::NP; ++p) {
            const int idx = tid + p * G::NT;
            if (idx < G::NA) {
                store_raw<16>(base + (idx / G::CPR) * G::ROWB + (idx % G::CPR) * 16, r[p]);
            } else if (idx < G::NA + G::NB) {
                const int j = idx - G::NA;
                if (!a.b_kn) {
                    store_raw<16>(base + G::A_BYTES + (j / G::CPR) * G::ROWB + (j % G::CPR) * 16, r[p]);
                } else {
                    const int kr = j / (G::TN / G::VE), jc = (j % (G::TN / G::VE)) * G::VE;
                    const in_t* e = reinterpret_cast<const in_t*>(&r[p]);
#pragma unroll
                    for (int q = 0; q < G::VE; ++q)
                        *reinterpret_cast<in_t*>(base + G::A_BYTES + (jc + q) * G::ROWB + kr * G::ELT) = e[q];
                }
            }
        }
    };

    accv_t acc[G::IM];
#pragma unroll
    for (int im = 0; im < G::IM; ++im)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[im][v] = 0;

    if (beg < end) {
        load(beg);
        write(0);
        __syncthreads();
        int buf = 0;
        for (int64_t q = beg; q < end; ++q) {
            const bool next = q + 1 < end;
            if (next) load(q + 1);
            if (wave_live) {
                const unsigned char* base = smem + buf * G::STAGE;
#pragma unroll
                for (int ks = 0; ks < BS / AT::KSTEP; ++ks) {
                    const int kk = (ks * AT::KSTEP + AT::KV * (lane >> 4)) * G::ELT;
                    const frag_t bf = bitcast<frag_t>(
                        load_raw<FB>(base + G::A_BYTES + (wave * 16 + (lane & 15)) * G::ROWB + kk));
#pragma unroll
                    for (int im = 0; im < G::IM; ++im) {
                        const frag_t af = bitcast<frag_t>(
                            load_raw<FB>(base + (im * 16 + (lane & 15)) * G::ROWB + kk));
#pragma unroll
                        for (int e = 0; e < AT::NSUB; ++e) AT::mma_sub(acc[im], af, bf, e);
                    }
                }
            }
            if (next) write(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }

    if (!wave_live) return;
    out_t* __restrict__ C = reinterpret_cast<out_t*>(a.C);
    const acc_t alpha = (acc_t)a.alpha, beta = (acc_t)a.beta;
    const int col = n0 + wave * 16 + (lane & 15);
#pragma unroll
    for (int im = 0; im < G::IM; ++im)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t row = (int64_t)i * BS + im * 16 + AT::out_row(lane, v);
            out_t* dst = C + row * a.ldc + col;
            acc_t val = alpha * acc[im][v];
            if (beta != (acc_t)0) val += beta * (acc_t)(*dst);
            *dst = (out_t)val;
        }
}

}  // namespace kami

using namespace kami;

extern "C" {

kami_status kami_bsr_spmm(const kami_bsr* A, const void* B, int32_t n, int64_t ldb,
                          int32_t b_layout, void* C, int64_t ldc, int32_t precision,
                          double alpha, double beta, void* stream) {
    if (!A || A->mb <= 0 || A->nb <= 0 || A->nnzb < 0 || !A->rowptr) return KAMI_EINVAL;
    if (A->nnzb > 0 && (!A->colidx || !A->vals)) return KAMI_EINVAL;
    if (b_layout != KAMI_B_NK && b_layout != KAMI_B_KN) return KAMI_EINVAL;
    if (n <= 0 || n % 16 || !B || !C) return KAMI_EINVAL;
    if (A->bs != 16 && A->bs != 32) return KAMI_ENOTSUP;
    int ei, eo;
    if (precision == KAMI_PREC_F16_F32 || precision == KAMI_PREC_BF16_F32) {
        ei = 2;
        eo = 4;
    } else if (precision == KAMI_PREC_F64) {
        ei = 8;
        eo = 8;
    } else {
        return KAMI_ENOTSUP;
    }
    const int64_t K = (int64_t)A->nb * A->bs;
    if (ldb == 0) ldb = b_layout == KAMI_B_NK ? K : n;
    if (ldc == 0) ldc = n;
    if ((b_layout == KAMI_B_NK && ldb < K) || (b_layout == KAMI_B_KN && ldb < n) || ldc < n)
        return KAMI_EINVAL;
    if ((ldb * ei) % 16 || ((uintptr_t)B % 16) || ((uintptr_t)C % eo)) return KAMI_EINVAL;
    if (A->nnzb > 0 && ((uintptr_t)A->vals % 16)) return KAMI_EINVAL;
    if ((n + 63) / 64 > 65535) return KAMI_ENOTSUP;
    kami_status st = device_check();
    if (st != KAMI_OK) return st;
    SpmmArgs a;
    a.mb = A->mb;
    a.nb = A->nb;
    a.n = n;
    a.b_kn = b_layout == KAMI_B_KN;
    a.nnzb = A->nnzb;
    a.ldb = ldb;
    a.ldc = ldc;
    a.rowptr = A->rowptr;
    a.colidx = A->colidx;
    a.vals = A->vals;
    a.B = B;
    a.C = C;
    a.alpha = alpha;
    a.beta = beta;
    const dim3 grid((unsigned)A->mb, (unsigned)((n + 63) / 64));
    hipStream_t s = (hipStream_t)stream;
#define KAMI_SPMM(BS, AT) hipLaunchKernelGGL((spmm_kernel<BS, AT>), grid, dim3(256), 0, s, a)
    if (precision == KAMI_PREC_F16_F32) {
        if (A->bs == 16) KAMI_SPMM(16, AtomF16x16); else KAMI_SPMM(32, AtomF16x32);
    } else if (precision == KAMI_PREC_BF16_F32) {
        if (A->bs == 16) KAMI_SPMM(16, AtomBF16x16); else KAMI_SPMM(32, AtomBF16x32);
    } else {
        if (A->bs == 16) KAMI_SPMM(16, AtomF64); else KAMI_SPMM(32, AtomF64);
    }
#undef KAMI_SPMM
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e);
    return KAMI_OK;
}

}  // extern "C"
