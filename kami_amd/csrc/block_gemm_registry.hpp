// Registry of compiled block-GEMM instances.
//
// The reference fixes shape, partition and wave count with -D macros per executable
// (scripts/fp64_block_square_GH200.sh:11-25, scripts/fp16_block_square_GH200.sh:11-25).
// Here every (precision, M, N, K, partition, P) instance is compiled once into the
// library and selected at run time from the descriptor.
#pragma once

#include <vector>

#include "block_gemm_kernel.hpp"
#include "block_gemm_pipe.hpp"

namespace kami {

typedef hipError_t (*block_launch_fn)(const BlockArgs&, int64_t, hipStream_t);

struct BlockInstance {
    int prec;       // kami_precision
    int m, n, k;
    int part;       // kami_partition
    int ranks;      // P as the reference counts it
    int pm, pn, pk; // wave grid
    int ks;         // K staged per LDS stage
    int threads;
    int lds;
    int atom_k;     // K of one fragment step
    block_launch_fn launch;
    int pipe;       // 1: one workgroup per problem; 2: persistent, deep prefetch
    int prefetch;   // K stages loaded ahead of the MFMAs
    int resident = 0;  // 1: one LDS stage holds all of K (chosen for iters > 1)
};

// Largest K stage (a divisor of BK, a multiple of PK * atom K step) whose LDS staging
// fits `budget` bytes, else the smallest such stage; whole-K staging uses one buffer,
// partial staging two.
template <class AT, int BM, int BN, int BK, int PM, int PN, int PK>
constexpr int choose_ks(int budget) {
    constexpr int elt = (int)sizeof(typename AT::in_t);
    int best = 0;
    for (int ks = BK; ks >= PK * AT::KSTEP; --ks) {
        if (BK % ks != 0 || ks % (PK * AT::KSTEP) != 0) continue;
        const int rowb = lds_row_bytes(ks * elt, AT::FRAG_BYTES);
        const int stage = ((PN == 1) ? 0 : BM * rowb) + BN * rowb;
        const int nbuf = (ks == BK) ? 1 : 2;
        best = ks;
        if (nbuf * stage <= budget) return ks;
    }
    return best;
}

constexpr int kStageBudget = 80 * 1024;

// KFENCE: see block_gemm_kernel (for stages of many K steps).
template <class AT, int PREC, int BM, int BN, int BK, int PM, int PN, int PK, int KFENCE = 0>
void add_instance(std::vector<BlockInstance>& v, int part, int ranks) {
    constexpr int KS = choose_ks<AT, BM, BN, BK, PM, PN, PK>(kStageBudget);
    using G = BlockGeom<AT, BM, BN, BK, PM, PN, PK, KS>;
    v.push_back(BlockInstance{PREC, BM, BN, BK, part, ranks, PM, PN, PK, KS, G::NT, G::LDS,
                              AT::KSTEP, &launch_block_gemm<AT, G, KFENCE>, 1, 1});
}

// A 3D instance whose full-width wave tiles would not fit the register file: the columns
// are computed in SP passes of BN / SP (the kernel's SP form), each a 3D product of its
// own with the depth reduction in LDS.
template <class AT, int PREC, int BM, int BN, int BK, int PM, int PN, int PK, int SP>
void add_pass_instance(std::vector<BlockInstance>& v, int ranks) {
    static_assert(BN % SP == 0, "column passes");
    constexpr int KS = choose_ks<AT, BM, BN / SP, BK, PM, PN, PK>(kStageBudget);
    using G = BlockGeom<AT, BM, BN / SP, BK, PM, PN, PK, KS>;
    v.push_back(BlockInstance{PREC, BM, BN, BK, 3, ranks, PM, PN, PK, KS, G::NT, G::LDS,
                              AT::KSTEP, &launch_block_gemm<AT, G, 0, SP>, 1, 1});
}

// Resident-repeat form of a 1D or 2D instance: one LDS stage holds the whole K of B (and
// of A in 2D; in 1D each wave keeps its rows of A in registers), so the repeated
// products of iters > 1 (the reference's NUM_ITER loop,
// src/block_gemm/KAMI/1d/fp64/block_gemm_1d_double_mma.cu:140) read nothing from memory
// after the first; the streaming form of the same key stages K in chunks instead.
// Chosen by the dispatcher when iters > 1.
template <class AT, int PREC, int BM, int BN, int BK, int PM, int PN = 1>
void add_resident_instance(std::vector<BlockInstance>& v, int ranks) {
    using G = BlockGeom<AT, BM, BN, BK, PM, PN, 1, BK>;
    static_assert(G::NS == 1, "whole K in one stage");
    v.push_back(BlockInstance{PREC, BM, BN, BK, PN == 1 ? 1 : 2, ranks, PM, PN, 1, BK, G::NT,
                              G::LDS, AT::KSTEP, &launch_block_gemm<AT, G, 2>, 1, 1, 1});
}

// Largest prefetch depth <= want that divides the stage count.
constexpr int pick_depth(int ns, int want) {
    for (int d = want; d > 1; --d)
        if (ns % d == 0) return d;
    return 1;
}

// Largest K stage whose two LDS buffers fit `budget` (the persistent form always
// double-buffers, also when one stage covers all of K).
template <class AT, int BM, int BN, int BK, int PM, int PN>
constexpr int choose_ks_pipe(int budget) {
    constexpr int elt = (int)sizeof(typename AT::in_t);
    int best = 0;
    for (int ks = BK; ks >= AT::KSTEP; --ks) {
        if (BK % ks != 0 || ks % AT::KSTEP != 0) continue;
        best = ks;
        if (2 * (BM + BN) * lds_row_bytes(ks * elt, AT::FRAG_BYTES) <= budget) return ks;
    }
    return best;
}

// Persistent 2D instance (PM x PN waves) with prefetch depth up to PD_WANT and K stages
// sized to an LDS budget of BUDGET bytes for the two buffers.
template <class AT, int PREC, int BM, int BN, int BK, int PM, int PN, int PD_WANT,
          int BUDGET = kStageBudget>
void add_pipe_instance(std::vector<BlockInstance>& v, int ranks) {
    constexpr int KS = choose_ks_pipe<AT, BM, BN, BK, PM, PN>(BUDGET);
    using G = BlockGeom<AT, BM, BN, BK, PM, PN, 1, KS>;
    constexpr int PD = pick_depth(G::NS, PD_WANT);
    v.push_back(BlockInstance{PREC, BM, BN, BK, 2, ranks, PM, PN, 1, KS, G::NT, 2 * G::STAGE_LDS,
                              AT::KSTEP, &launch_block_gemm_pipe<AT, G, PD>, 2, PD});
}

void register_f64_instances(std::vector<BlockInstance>& v);
void register_f16_instances(std::vector<BlockInstance>& v);
void register_bf16_instances(std::vector<BlockInstance>& v);

}  // namespace kami
