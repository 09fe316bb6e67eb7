// Block SpMM: C (M x N) = alpha * A * op(B) + beta * C with A in BSR and B dense.
//
// The reference computes one such product per CTA with A re-tiled per warp rank
// (src/sparse/spmm/2d/block_spmm_2d_half_mma.cu:72-235, bsr2tilebsr :264-291) and B stored
// N x K, so C(i, j) = sum_k A(i, k) B[j * K + k] (check loop :378-403).  Here A stays in
// plain BSR (the same format the SpGEMM entry points take) and the product is spread
// over workgroups:
//   * one workgroup per (block row i, 64-column tile of C).  FP16 / BF16 (split form,
//     spmm_split_kernel): the row's blocks are shared out over the four waves, each wave
//     computing all 64 columns of its blocks from fragments loaded straight from memory,
//     and the four partial tiles are summed through LDS in a fixed order.  FP64
//     (spmm_kernel): wave w owns the 16 columns [16 w, 16 w + 16) of the tile and all of
//     the row's blocks, as below;
//   * only the A block is shared by the four waves, so only it is staged in LDS (rows
//     padded to the conflict-free stride, double-buffered, one barrier per block); each
//     lane loads its own B values straight into MFMA fragments, in the fragment layout of
//     the MFMA atom (for column c, k-step t and lane group g the KV values from
//     t KSTEP + g KV), so one load instruction reads four consecutive fragments (64 bytes
//     with 16-byte fragments) of every column it touches;
//   * the next block's A piece and B fragments are loaded while the current block is
//     multiplied; the column indices of a row are read 64 at a time into a register and
//     taken out with readlane, so no scalar load waits in the loop;
//   * workgroups are numbered so that each XCD (workgroup L runs on XCD L mod 8) takes a
//     contiguous run of (column tile, block row) pairs: its L2 sees the B columns of few
//     tiles;
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
    static constexpr int FB = AT::FRAG_BYTES;
    static constexpr int ROWB = lds_row_bytes(BS * ELT, FB);
    static constexpr int STAGE = BS * ROWB;  // one A block
    static constexpr int CPR = BS / VE;      // 16-byte pieces per A row
    static constexpr int NA = BS * CPR;      // pieces of an A block
    static constexpr int NPA = (NA + NT - 1) / NT;
    static constexpr int IM = BS / 16;
    static constexpr int NKS = BS / AT::KSTEP;  // MFMA k-steps per block
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
    constexpr int FB = AT::FRAG_BYTES, KV = AT::KV;
    struct elems_t {
        in_t v[KV];
    };
    __shared__ __attribute__((aligned(16))) unsigned char smem[2 * G::STAGE];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, grp = lane >> 4;
    // XCD x = L mod 8 takes the contiguous run [x q + min(x, r), ...) of the tile-major
    // order, q = total / 8, r = total mod 8 (a bijection of [0, total)).
    const uint32_t L = blockIdx.x, total = gridDim.x;
    const uint32_t x = L % 8, q8 = total / 8, r8 = total % 8;
    const uint32_t P = x * q8 + (x < r8 ? x : r8) + L / 8;
    const int i = (int)(P % (uint32_t)a.mb);      // block row of A and C
    const int n0 = (int)(P / (uint32_t)a.mb) * G::TN;  // first column of the C tile
    const int col = n0 + wave * 16 + (lane & 15);
    const bool wave_live = n0 + wave * 16 < a.n;  // wave-uniform; n is a multiple of 16
    const in_t* __restrict__ vals = reinterpret_cast<const in_t*>(a.vals);
    const in_t* __restrict__ B = reinterpret_cast<const in_t*>(a.B);
    // stored blocks of row i (clamped to the arrays the caller described)
    const int64_t rb = a.rowptr[i], re = a.rowptr[i + 1];
    const int64_t beg = rb < 0 ? 0 : rb;
    const int64_t end = re > a.nnzb ? a.nnzb : re;

    // column indices [cbase, cbase + 64) of the row, one per lane
    int64_t cbase = beg;
    int cvec = beg + lane < end ? a.colidx[beg + lane] : -1;
    auto block_col = [&](int64_t q) {  // q uniform, called with q = beg, beg + 1, ...
        if (q - cbase >= 64) {
            cbase += 64;
            cvec = cbase + lane < end ? a.colidx[cbase + lane] : -1;
        }
        return __builtin_amdgcn_readlane(cvec, (int)(q - cbase));
    };

    u4_t ra[G::NPA];
    auto load_a = [&](int64_t q) {
#pragma unroll
        for (int p = 0; p < G::NPA; ++p) {
            const int idx = tid + p * G::NT;
            if (idx < G::NA)
                ra[p] = load_raw<16>(vals + q * BS * BS + (idx / G::CPR) * BS + (idx % G::CPR) * G::VE);
        }
    };
    auto write_a = [&](int buf) {
        unsigned char* base = smem + buf * G::STAGE;
#pragma unroll
        for (int p = 0; p < G::NPA; ++p) {
            const int idx = tid + p * G::NT;
            if (idx < G::NA)
                store_raw<16>(base + (idx / G::CPR) * G::ROWB + (idx % G::CPR) * 16, ra[p]);
        }
    };
    // B fragments of block column kb for this lane: step t holds k = kb bs + t KSTEP + g KV + v
    auto load_b = [&](int kb, frag_t (&f)[G::NKS]) {
        elems_t z;
#pragma unroll
        for (int v = 0; v < KV; ++v) z.v[v] = in_t(0);
#pragma unroll
        for (int t = 0; t < G::NKS; ++t) f[t] = bitcast<frag_t>(z);
        if (!wave_live || kb < 0 || kb >= a.nb) return;  // wave-uniform
        const int64_t k0 = (int64_t)kb * BS + grp * KV;
        if (!a.b_kn) {
            const in_t* p = B + (int64_t)col * a.ldb + k0;
#pragma unroll
            for (int t = 0; t < G::NKS; ++t) f[t] = bitcast<frag_t>(load_raw<FB>(p + t * AT::KSTEP));
        } else {
            const in_t* p = B + k0 * a.ldb + col;
#pragma unroll
            for (int t = 0; t < G::NKS; ++t) {
                elems_t e;
#pragma unroll
                for (int v = 0; v < KV; ++v) e.v[v] = p[(int64_t)(t * AT::KSTEP + v) * a.ldb];
                f[t] = bitcast<frag_t>(e);
            }
        }
    };

    accv_t acc[G::IM];
#pragma unroll
    for (int im = 0; im < G::IM; ++im)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[im][v] = 0;

    if (beg < end) {
        frag_t bc[G::NKS], bn[G::NKS];
        load_a(beg);
        load_b(block_col(beg), bc);
        write_a(0);
        __syncthreads();
        int buf = 0;
        for (int64_t q = beg; q < end; ++q) {
            const bool next = q + 1 < end;
            if (next) {
                load_a(q + 1);
                load_b(block_col(q + 1), bn);
            }
            if (wave_live) {
                const unsigned char* base = smem + buf * G::STAGE;
#pragma unroll
                for (int ks = 0; ks < G::NKS; ++ks) {
                    const int kk = (ks * AT::KSTEP + KV * grp) * G::ELT;
                    frag_t af[G::IM];
#pragma unroll
                    for (int im = 0; im < G::IM; ++im)
                        af[im] = bitcast<frag_t>(load_raw<FB>(base + (im * 16 + (lane & 15)) * G::ROWB + kk));
#pragma unroll
                    for (int e = 0; e < AT::NSUB; ++e)
#pragma unroll
                        for (int im = 0; im < G::IM; ++im) AT::mma_sub(acc[im], af[im], bc[ks], e);
                }
            }
            if (next) write_a(buf ^ 1);
            __syncthreads();
            buf ^= 1;
#pragma unroll
            for (int t = 0; t < G::NKS; ++t) bc[t] = bn[t];
        }
    }

    if (!wave_live) return;
    out_t* __restrict__ C = reinterpret_cast<out_t*>(a.C);
    const acc_t alpha = (acc_t)a.alpha, beta = (acc_t)a.beta;
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

// ---- split form: the blocks of a row shared by the workgroup's four waves ----
//
// One workgroup per (block row i, tile of TN = 16 CT columns of C).  Wave w takes the
// row's blocks w, w + 4, w + 8, ..., so the longest row's chain of load-multiply steps
// is a quarter as long; each wave loads the A block's and B's MFMA fragments straight
// from memory (no LDS staging, no barrier in the loop), the next block's while the
// current one is multiplied.  The four partial tiles are summed through LDS in wave
// order 0, 1, 2, 3 (the same order every run), then C = alpha sum + beta C.
template <int BS, class AT, int CT>
struct SpmmSplitGeom {
    using in_t = typename AT::in_t;
    using acc_t = typename AT::acc_t;
    static constexpr int NT = 256, WAVES = 4, TN = 16 * CT;
    static constexpr int IM = BS / 16;
    static constexpr int NKS = BS / AT::KSTEP;  // MFMA k-steps per block
    static constexpr int NE = IM * CT * 4;      // accumulator values per lane
    static constexpr int LDS = WAVES * NE * 64 * (int)sizeof(acc_t);
    static_assert(BS % AT::KSTEP == 0 && BS % 16 == 0, "block size");
    static_assert(LDS <= 64 * 1024, "partial tiles");
};

template <int BS, class AT, int CT>
__global__ void __launch_bounds__(256) spmm_split_kernel(SpmmArgs a) {
    using G = SpmmSplitGeom<BS, AT, CT>;
    using in_t = typename AT::in_t;
    using acc_t = typename AT::acc_t;
    using out_t = typename AT::out_t;
    using frag_t = typename AT::frag_t;
    using accv_t = typename AT::accv_t;
    constexpr int FB = AT::FRAG_BYTES, KV = AT::KV;
    struct elems_t {
        in_t v[KV];
    };
    __shared__ __attribute__((aligned(16))) acc_t part[G::WAVES * G::NE * 64];

    const int tid = threadIdx.x, lane = tid & 63, grp = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // XCD x = L mod 8 takes the contiguous run [x q + min(x, r), ...) of the tile-major
    // order, q = total / 8, r = total mod 8 (a bijection of [0, total)).
    const uint32_t L = blockIdx.x, total = gridDim.x;
    const uint32_t x = L % 8, q8 = total / 8, r8 = total % 8;
    const uint32_t P = x * q8 + (x < r8 ? x : r8) + L / 8;
    const int i = (int)(P % (uint32_t)a.mb);           // block row of A and C
    const int n0 = (int)(P / (uint32_t)a.mb) * G::TN;  // first column of the C tile
    const in_t* __restrict__ vals = reinterpret_cast<const in_t*>(a.vals);
    const in_t* __restrict__ B = reinterpret_cast<const in_t*>(a.B);
    const int64_t rb = a.rowptr[i], re = a.rowptr[i + 1];
    const int64_t beg = rb < 0 ? 0 : rb;
    const int64_t end = re > a.nnzb ? a.nnzb : re;

    // the row's column indices [cbase, cbase + 64), one per lane, taken out with readlane
    // (q grows with each call of a wave, so the window only moves forward)
    int64_t cbase = beg;
    int cvec = beg + lane < end ? a.colidx[beg + lane] : -1;
    auto block_col = [&](int64_t q) {
        if (q - cbase >= 64) {
            cbase += (q - cbase) / 64 * 64;
            cvec = cbase + lane < end ? a.colidx[cbase + lane] : -1;
        }
        return __builtin_amdgcn_readlane(cvec, (int)(q - cbase));
    };
    // fragments of one block: A (IM atom rows x NKS k-steps) and B (CT atoms x NKS)
    struct Frags {
        frag_t a[G::IM][G::NKS];
        frag_t b[CT][G::NKS];
    };
    auto load = [&](int64_t q, Frags& f) {
        const int kb = block_col(q);
        const in_t* ab = vals + q * BS * BS + (lane & 15) * BS + grp * KV;
#pragma unroll
        for (int im = 0; im < G::IM; ++im)
#pragma unroll
            for (int t = 0; t < G::NKS; ++t)
                f.a[im][t] = bitcast<frag_t>(load_raw<FB>(ab + im * 16 * BS + t * AT::KSTEP));
        const int64_t k0 = (int64_t)kb * BS + grp * KV;
        const bool kb_ok = kb >= 0 && kb < a.nb;  // wave-uniform
#pragma unroll
        for (int jn = 0; jn < CT; ++jn) {
            const int col = n0 + jn * 16 + (lane & 15);
            const bool live = kb_ok && n0 + jn * 16 < a.n;  // wave-uniform (n % 16 == 0)
#pragma unroll
            for (int t = 0; t < G::NKS; ++t) {
                elems_t e;
                if (!live) {
#pragma unroll
                    for (int v = 0; v < KV; ++v) e.v[v] = in_t(0);
                    f.b[jn][t] = bitcast<frag_t>(e);
                } else if (!a.b_kn) {
                    f.b[jn][t] = bitcast<frag_t>(
                        load_raw<FB>(B + (int64_t)col * a.ldb + k0 + t * AT::KSTEP));
                } else {
                    const in_t* p = B + k0 * a.ldb + col;
#pragma unroll
                    for (int v = 0; v < KV; ++v) e.v[v] = p[(int64_t)(t * AT::KSTEP + v) * a.ldb];
                    f.b[jn][t] = bitcast<frag_t>(e);
                }
            }
        }
    };

    accv_t acc[G::IM][CT];
#pragma unroll
    for (int im = 0; im < G::IM; ++im)
#pragma unroll
        for (int jn = 0; jn < CT; ++jn)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[im][jn][v] = 0;

    int64_t q = beg + wave;
    if (q < end) {
        Frags cur, nxt;
        load(q, cur);
        for (; q < end; q += G::WAVES) {
            const bool more = q + G::WAVES < end;  // wave-uniform
            if (more) load(q + G::WAVES, nxt);
#pragma unroll
            for (int t = 0; t < G::NKS; ++t)
#pragma unroll
                for (int e = 0; e < AT::NSUB; ++e)
#pragma unroll
                    for (int jn = 0; jn < CT; ++jn)
#pragma unroll
                        for (int im = 0; im < G::IM; ++im) AT::mma_sub(acc[im][jn], cur.a[im][t], cur.b[jn][t], e);
            if (more) cur = nxt;
        }
    }

    // partial tiles: value (im, jn, v) of lane l of wave w at ((w NE + (im CT + jn) 4 + v) 64 + l)
#pragma unroll
    for (int im = 0; im < G::IM; ++im)
#pragma unroll
        for (int jn = 0; jn < CT; ++jn)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                part[((wave * G::NE + (im * CT + jn) * 4 + v) << 6) + lane] = acc[im][jn][v];
    __syncthreads();
    out_t* __restrict__ C = reinterpret_cast<out_t*>(a.C);
    const acc_t alpha = (acc_t)a.alpha, beta = (acc_t)a.beta;
#pragma unroll
    for (int s = 0; s < G::NE / 4; ++s) {
        const int e = tid + G::NT * s;  // element (im, jn, v) = e / 64, lane e % 64
        acc_t sum = part[e];
#pragma unroll
        for (int w = 1; w < G::WAVES; ++w) sum += part[w * G::NE * 64 + e];
        const int iv = e >> 6, l = e & 63;
        const int im = iv / (CT * 4), jn = (iv / 4) % CT, v = iv % 4;
        const int col = n0 + jn * 16 + (l & 15);
        if (col >= a.n) continue;
        const int64_t row = (int64_t)i * BS + im * 16 + AT::out_row(l, v);
        out_t* dst = C + row * a.ldc + col;
        acc_t val = alpha * sum;
        if (beta != (acc_t)0) val += beta * (acc_t)(*dst);
        *dst = (out_t)val;
    }
}

// The split form serves FP16 / BF16 (64-column tiles, 88 VGPRs: every workgroup of the
// w6 product resident at once; kernel 10.6 -> 9.2 µs).  FP64 keeps the one-wave-per-16-
// columns form: its split tiles need 198 VGPRs (two waves per SIMD), and the 2048
// workgroups then run in four rounds (33.7 -> 39.8 µs); with 16-column split tiles (144
// VGPRs, every column tile loading the row's A blocks again) it ran 42.3-42.6 µs.
constexpr int kSpmmSplitAtoms = 4;

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
    if (!span_fits(b_layout == KAMI_B_NK ? n : K, ldb, ei) ||
        !span_fits((int64_t)A->mb * A->bs, ldc, eo))
        return KAMI_EINVAL;
    if ((ldb * ei) % 16 || ((uintptr_t)B % 16) || ((uintptr_t)C % eo)) return KAMI_EINVAL;
    if (A->nnzb > 0 && ((uintptr_t)A->vals % 16)) return KAMI_EINVAL;
    // workgroups: one per (block row, 64-column tile of C) in both forms
    const int64_t wgs = (int64_t)A->mb * (((int64_t)n + 63) / 64);
    if (wgs >= ((int64_t)1 << 24)) return KAMI_ENOTSUP;  // work-items of one dispatch < 2^32
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
    const dim3 grid((unsigned)wgs);
    hipStream_t s = (hipStream_t)stream;
#define KAMI_SPMM_SPLIT(BS, AT) \
    hipLaunchKernelGGL((spmm_split_kernel<BS, AT, kSpmmSplitAtoms>), grid, dim3(256), 0, s, a)
#define KAMI_SPMM(BS, AT) hipLaunchKernelGGL((spmm_kernel<BS, AT>), grid, dim3(256), 0, s, a)
    if (precision == KAMI_PREC_F16_F32) {
        if (A->bs == 16) KAMI_SPMM_SPLIT(16, AtomF16x16); else KAMI_SPMM_SPLIT(32, AtomF16x32);
    } else if (precision == KAMI_PREC_BF16_F32) {
        if (A->bs == 16) KAMI_SPMM_SPLIT(16, AtomBF16x16); else KAMI_SPMM_SPLIT(32, AtomBF16x32);
    } else {
        if (A->bs == 16) KAMI_SPMM(16, AtomF64); else KAMI_SPMM(32, AtomF64);
    }
#undef KAMI_SPMM
#undef KAMI_SPMM_SPLIT
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e);
    return KAMI_OK;
}

}  // extern "C"
