// Block-level GEMM: one problem per workgroup, waves on a PM x PN x PK grid.
//
// This is the MI355X form of KAMI's communication-avoiding block GEMM
// (src/block_gemm/KAMI/{1d,2d,3d}/fp64/block_gemm_{1d,2d,3d}_double_mma.cu and the
// fp16 variants).  The partition names keep the reference's meaning:
//   1D  (PM = P, PN = 1, PK = 1): each wave owns M/P rows of A, held in registers for
//       the whole product (the reference's a_frags, F64-1D:73-102); B is the only
//       operand that crosses waves and it goes through LDS (F64-1D:148-185).
//   2D  (PM = PN = P, PK = 1): wave (r, c) owns C(r, c); A rows are shared along a wave
//       row and B columns along a wave column through LDS (F64-2D:156-186).
//   3D  (PK > 1): as 2D, and K is split over a depth axis; the PK partial C tiles are
//       reduced on the device through LDS (the reference sums slabs on the host,
//       F64-3D:324-333).
// Differences from the reference that are deliberate:
//   * operands shared between waves are staged into LDS once per K stage with 16-byte
//     coalesced loads, not written by an owner wave from its registers, and the stage
//     buffers are double-buffered so one barrier per stage suffices (the reference's
//     2D kernel is missing its post-MMA barrier, F64-2D:188-212);
//   * the wave64 MFMA atoms of kami_device.hpp replace mma.sync m16n8k8 / m16n8k16;
//   * alpha and beta are honoured (the reference ignores them, F64-1D:55 vs :205-208);
//   * `iters` repeats the whole product into the same accumulators, which is the
//     reference's resident-repeat NUM_ITER loop (F64-1D:140); iters = 1 is a plain GEMM.
#pragma once

#include "kami_device.hpp"

namespace kami {

constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <class AT, int BM_, int BN_, int BK_, int PM_, int PN_, int PK_, int KS_>
struct BlockGeom {
    static constexpr int BM = BM_, BN = BN_, BK = BK_;
    static constexpr int PM = PM_, PN = PN_, PK = PK_, KS = KS_;
    static constexpr int WAVES = PM * PN * PK;
    static constexpr int NT = WAVES * kWave;
    static constexpr int TM = BM / PM, TN = BN / PN;
    // atom edge (16, or 32 for the 32 x 32 atoms) and accumulators per lane
    static constexpr int DIM = AtomShape<AT>::DIM, NACC = AtomShape<AT>::NACC;
    static constexpr int IM = TM / DIM, JN = TN / DIM;
    static constexpr int NS = BK / KS;
    static constexpr int KSW = KS / PK;
    static constexpr int KSTEPS = KSW / AT::KSTEP;
    static constexpr int ELT = (int)sizeof(typename AT::in_t);
    static constexpr int VE = 16 / ELT;  // elements per 16-byte staging chunk
    static constexpr bool A_PRIV = (PN == 1);
    static constexpr int NBUF = (NS == 1) ? 1 : 2;
    static constexpr int ROWB = lds_row_bytes(KS * ELT, AT::FRAG_BYTES);  // conflict-free reads
    static constexpr int A_LDS = A_PRIV ? 0 : BM * ROWB;
    static constexpr int B_LDS = BN * ROWB;
    static constexpr int STAGE_LDS = A_LDS + B_LDS;
    static constexpr int STAGING = NBUF * STAGE_LDS;
    static constexpr int ACC_ELT = (int)sizeof(typename AT::acc_t);
    static constexpr int RED_ROW_BYTES = PK * PM * PN * JN * NACC * 64 * ACC_ELT;
    static constexpr int RED_ROWS = (PK > 1) ? cmin(IM, cmax(1, STAGING / RED_ROW_BYTES)) : 1;
    static constexpr int RED_ROUNDS = (IM + RED_ROWS - 1) / RED_ROWS;
    static constexpr int RED_LDS = (PK > 1) ? RED_ROWS * RED_ROW_BYTES : 0;
    static constexpr int LDS = cmax(STAGING, RED_LDS);
    static constexpr int CA = A_PRIV ? 0 : BM * KS / VE;  // 16-byte chunks of an A stage
    static constexpr int CB = BN * KS / VE;               // 16-byte chunks of a B stage
    static constexpr int NA = (CA + NT - 1) / NT;
    static constexpr int NBC = (CB + NT - 1) / NT;

    static_assert(BM % PM == 0 && BN % PN == 0 && BK % KS == 0 && KS % PK == 0, "tiling");
    static_assert(TM % DIM == 0 && TN % DIM == 0, "wave tile must be a multiple of the atom");
    static_assert(KSW % AT::KSTEP == 0, "wave K slice must be a multiple of the atom K step");
    static_assert((KS * ELT) % 16 == 0 && (BN * ELT) % 16 == 0, "16-byte staging");
    static_assert(NT <= 1024, "workgroup too large");
    static_assert(LDS <= 160 * 1024, "LDS budget of one workgroup");
};

struct BlockArgs {
    const void* A;
    const void* B;
    void* C;
    int64_t sa, sb, sc;       // per-problem strides (elements)
    int64_t lda, ldb, ldc;    // leading dimensions (elements)
    double alpha, beta;
    int32_t iters;
    int32_t b_kn;             // 1: B stored K x N
};

// KFENCE > 0: a scheduling fence after every KFENCE K steps of a stage, so the
// compiler does not hoist the B fragment loads of a whole long stage ahead of the MFMAs
// (used by the resident forms, whose single stage covers all of K).
//
// SP > 1: the wave tiles of the full BN columns do not fit the register file, so the
// problem's columns are computed in SP passes of G::BN = BN / SP columns each (B rows and
// C columns of pass sp start at sp * G::BN); A is read again in every pass.
template <class AT, class G, int KFENCE = 0, int SP = 1>
__global__ void __launch_bounds__(G::NT)
block_gemm_kernel(BlockArgs args) {
    using in_t = typename AT::in_t;
    using acc_t = typename AT::acc_t;
    using out_t = typename AT::out_t;
    using frag_t = typename AT::frag_t;
    using accv_t = typename AT::accv_t;
    constexpr int FB = AT::FRAG_BYTES;

    __shared__ __attribute__((aligned(16))) unsigned char smem[G::LDS];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    // wave grid coordinates: depth fastest, then column, then row
    const int pk = wave % G::PK;
    const int pn = (wave / G::PK) % G::PN;
    const int pm = wave / (G::PK * G::PN);

    const int64_t prob = blockIdx.x;
    const int64_t lda = args.lda, ldb = args.ldb, ldc = args.ldc;
    const bool b_kn = args.b_kn != 0;
    const in_t* __restrict__ A = reinterpret_cast<const in_t*>(args.A) + prob * args.sa;
#pragma unroll 1
    for (int sp = 0; sp < SP; ++sp) {
    if (sp > 0) __syncthreads();  // the previous pass is done with LDS
    const in_t* __restrict__ B = reinterpret_cast<const in_t*>(args.B) + prob * args.sb +
                                 (int64_t)sp * G::BN * (b_kn ? 1 : ldb);
    out_t* __restrict__ C = reinterpret_cast<out_t*>(args.C) + prob * args.sc + sp * G::BN;

    accv_t acc[G::IM][G::JN];
#pragma unroll
    for (int i = 0; i < G::IM; ++i)
#pragma unroll
        for (int j = 0; j < G::JN; ++j)
#pragma unroll
            for (int v = 0; v < G::NACC; ++v) acc[i][j][v] = 0;

    // ---- staging of the shared operands (global -> registers -> LDS) ----
    u4_t stA[G::NA > 0 ? G::NA : 1];
    u4_t stB[G::NBC];
    auto load_stage = [&](int s) {
        const int k0 = s * G::KS;
        if constexpr (!G::A_PRIV) {
            constexpr int CPR = G::KS / G::VE;
#pragma unroll
            for (int i = 0; i < G::NA; ++i) {
                const int c = tid + i * G::NT;
                if (c < G::CA) {
                    const int r = c / CPR, cc = c % CPR;
                    stA[i] = load_raw<16>(A + r * lda + k0 + cc * G::VE);
                }
            }
        }
        if (!b_kn) {
            constexpr int CPR = G::KS / G::VE;
#pragma unroll
            for (int i = 0; i < G::NBC; ++i) {
                const int c = tid + i * G::NT;
                if (c < G::CB) {
                    const int r = c / CPR, cc = c % CPR;
                    stB[i] = load_raw<16>(B + r * ldb + k0 + cc * G::VE);
                }
            }
        } else {
            constexpr int CPR = G::BN / G::VE;
#pragma unroll
            for (int i = 0; i < G::NBC; ++i) {
                const int c = tid + i * G::NT;
                if (c < G::CB) {
                    const int kr = c / CPR, jc = c % CPR;
                    stB[i] = load_raw<16>(B + (int64_t)(k0 + kr) * ldb + jc * G::VE);
                }
            }
        }
    };
    auto write_stage = [&](int buf) {
        unsigned char* base = smem + buf * G::STAGE_LDS;
        if constexpr (!G::A_PRIV) {
            constexpr int CPR = G::KS / G::VE;
#pragma unroll
            for (int i = 0; i < G::NA; ++i) {
                const int c = tid + i * G::NT;
                if (c < G::CA) {
                    const int r = c / CPR, cc = c % CPR;
                    store_raw<16>(base + r * G::ROWB + cc * 16, stA[i]);
                }
            }
        }
        unsigned char* bb = base + G::A_LDS;
        if (!b_kn) {
            constexpr int CPR = G::KS / G::VE;
#pragma unroll
            for (int i = 0; i < G::NBC; ++i) {
                const int c = tid + i * G::NT;
                if (c < G::CB) {
                    const int r = c / CPR, cc = c % CPR;
                    store_raw<16>(bb + r * G::ROWB + cc * 16, stB[i]);
                }
            }
        } else {
            // transpose K x N chunks into the N-major LDS layout
            constexpr int CPR = G::BN / G::VE;
#pragma unroll
            for (int i = 0; i < G::NBC; ++i) {
                const int c = tid + i * G::NT;
                if (c < G::CB) {
                    const int kr = c / CPR, jc = c % CPR;
                    const in_t* e = reinterpret_cast<const in_t*>(&stB[i]);
#pragma unroll
                    for (int q = 0; q < G::VE; ++q)
                        *reinterpret_cast<in_t*>(bb + (jc * G::VE + q) * G::ROWB + kr * G::ELT) = e[q];
                }
            }
        }
    };

    // ---- register-resident A (1D): each wave loads its own rows in fragment layout ----
    constexpr int AIM = G::A_PRIV ? G::IM : 1;
    constexpr int AKS = G::A_PRIV ? G::KSTEPS : 1;
    frag_t a_cur[AIM][AKS];
    frag_t a_nxt[(G::A_PRIV && G::NS > 1) ? AIM : 1][(G::A_PRIV && G::NS > 1) ? AKS : 1];
    auto load_a = [&](int s, frag_t (&dst)[AIM][AKS]) {
        if constexpr (G::A_PRIV) {
            const int kb = s * G::KS + pk * G::KSW + AT::KV * (lane / G::DIM);
#pragma unroll
            for (int im = 0; im < G::IM; ++im) {
                const in_t* row = A + (int64_t)(pm * G::TM + im * G::DIM + (lane % G::DIM)) * lda + kb;
#pragma unroll
                for (int ks = 0; ks < G::KSTEPS; ++ks)
                    dst[im][ks] = bitcast<frag_t>(load_raw<FB>(row + ks * AT::KSTEP));
            }
        }
    };

    auto compute = [&](int buf) {
        const unsigned char* base = smem + buf * G::STAGE_LDS;
        const unsigned char* bb = base + G::A_LDS;
        const int kin = pk * G::KSW + AT::KV * (lane / G::DIM);
#pragma unroll
        for (int ks = 0; ks < G::KSTEPS; ++ks) {
            const int kk = (kin + ks * AT::KSTEP) * G::ELT;
            // A fragments of this K step stay live; B fragments are streamed one column
            // atom at a time so the accumulators keep most of the register file.
            frag_t af[G::IM];
#pragma unroll
            for (int im = 0; im < G::IM; ++im) {
                if constexpr (G::A_PRIV) {
                    af[im] = a_cur[im][ks];
                } else {
                    af[im] = bitcast<frag_t>(
                        load_raw<FB>(base + (pm * G::TM + im * G::DIM + (lane % G::DIM)) * G::ROWB + kk));
                }
            }
            if constexpr (AT::NSUB == 1) {
#pragma unroll
                for (int jn = 0; jn < G::JN; ++jn) {
                    const frag_t bf = bitcast<frag_t>(
                        load_raw<FB>(bb + (pn * G::TN + jn * G::DIM + (lane % G::DIM)) * G::ROWB + kk));
#pragma unroll
                    for (int im = 0; im < G::IM; ++im) AT::mma(acc[im][jn], af[im], bf);
                }
            } else {
                // several instructions per fragment: issue sub-step e for every atom
                // before sub-step e + 1 so consecutive MFMAs never share an accumulator
                frag_t bf[G::JN];
#pragma unroll
                for (int jn = 0; jn < G::JN; ++jn)
                    bf[jn] = bitcast<frag_t>(
                        load_raw<FB>(bb + (pn * G::TN + jn * G::DIM + (lane % G::DIM)) * G::ROWB + kk));
#pragma unroll
                for (int e = 0; e < AT::NSUB; ++e)
#pragma unroll
                    for (int jn = 0; jn < G::JN; ++jn)
#pragma unroll
                        for (int im = 0; im < G::IM; ++im) AT::mma_sub(acc[im][jn], af[im], bf[jn], e);
            }
            if constexpr (KFENCE > 0)
                if ((ks + 1) % KFENCE == 0) __builtin_amdgcn_sched_barrier(0);
        }
    };

    // ---- main loop ----
    load_stage(0);
    load_a(0, a_cur);
    write_stage(0);
    __syncthreads();
    if constexpr (G::NS == 1) {
        for (int it = 0; it < args.iters; ++it) compute(0);
    } else {
        int g = 0;
        for (int it = 0; it < args.iters; ++it) {
#pragma unroll 1
            for (int s = 0; s < G::NS; ++s, ++g) {
                const bool has_next = (s + 1 < G::NS) || (it + 1 < args.iters);
                const int sn = (s + 1 == G::NS) ? 0 : s + 1;
                if (has_next) {
                    load_stage(sn);
                    if constexpr (G::A_PRIV) load_a(sn, a_nxt);
                }
                compute(g & 1);
                if (has_next) {
                    write_stage((g & 1) ^ 1);
                    if constexpr (G::A_PRIV) {
#pragma unroll
                        for (int im = 0; im < AIM; ++im)
#pragma unroll
                            for (int ks = 0; ks < AKS; ++ks) a_cur[im][ks] = a_nxt[im][ks];
                    }
                }
                __syncthreads();
            }
        }
    }

    // ---- 3D: reduce the PK partial tiles through LDS; atom (im, jn) is finalised by
    //      depth (im * JN + jn) % PK, summing depths in ascending order ----
    if constexpr (G::PK > 1) {
        acc_t* red = reinterpret_cast<acc_t*>(smem);
#pragma unroll
        for (int r = 0; r < G::RED_ROUNDS; ++r) {
            __syncthreads();
#pragma unroll
            for (int im = r * G::RED_ROWS; im < cmin((r + 1) * G::RED_ROWS, G::IM); ++im)
#pragma unroll
                for (int jn = 0; jn < G::JN; ++jn) {
                    if ((im * G::JN + jn) % G::PK != pk) {
                        const int slot =
                            (((pk * G::PM + pm) * G::PN + pn) * G::RED_ROWS + (im - r * G::RED_ROWS)) *
                                G::JN + jn;
#pragma unroll
                        for (int v = 0; v < G::NACC; ++v)
                            red[slot * (G::NACC * 64) + v * 64 + lane] = acc[im][jn][v];
                    }
                }
            __syncthreads();
#pragma unroll
            for (int im = r * G::RED_ROWS; im < cmin((r + 1) * G::RED_ROWS, G::IM); ++im)
#pragma unroll
                for (int jn = 0; jn < G::JN; ++jn) {
                    if ((im * G::JN + jn) % G::PK == pk) {
                        accv_t tot;
#pragma unroll
                        for (int q = 0; q < G::PK; ++q) {
                            accv_t x;
                            if (q == pk) {
                                x = acc[im][jn];
                            } else {
                                const int slot =
                                    (((q * G::PM + pm) * G::PN + pn) * G::RED_ROWS +
                                     (im - r * G::RED_ROWS)) * G::JN + jn;
#pragma unroll
                                for (int v = 0; v < G::NACC; ++v)
                                    x[v] = red[slot * (G::NACC * 64) + v * 64 + lane];
                            }
                            if (q == 0) tot = x; else tot += x;
                        }
                        acc[im][jn] = tot;
                    }
                }
        }
    }

    // ---- epilogue: C = alpha * acc + beta * C ----
    const acc_t alpha = (acc_t)args.alpha;
    const acc_t beta = (acc_t)args.beta;
#pragma unroll
    for (int im = 0; im < G::IM; ++im)
#pragma unroll
        for (int jn = 0; jn < G::JN; ++jn) {
            if (G::PK > 1 && (im * G::JN + jn) % G::PK != pk) continue;
            const int col = pn * G::TN + jn * G::DIM + (lane % G::DIM);
#pragma unroll
            for (int v = 0; v < G::NACC; ++v) {
                const int row = pm * G::TM + im * G::DIM + AT::out_row(lane, v);
                out_t* dst = C + (int64_t)row * ldc + col;
                acc_t val = alpha * acc[im][jn][v];
                if (beta != (acc_t)0) val += beta * (acc_t)(*dst);
                *dst = (out_t)val;
            }
        }
    }  // column passes
}

template <class AT, class G, int KFENCE = 0, int SP = 1>
hipError_t launch_block_gemm(const BlockArgs& args, int64_t count, hipStream_t stream) {
    hipLaunchKernelGGL((block_gemm_kernel<AT, G, KFENCE, SP>), dim3((unsigned)count), dim3(G::NT),
                       0, stream, args);
    return hipGetLastError();
}

}  // namespace kami
