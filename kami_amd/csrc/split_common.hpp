// Types shared by the split-precision FP64 translation units (FP16/BF16 digits in
// split_f64.hip, int8 digits in split_i8.hip): kernel arguments and the registry entry
// that the descriptor dispatch of kami_block_gemm / kami_split_gemm resolves.
#pragma once

#include <cstdint>
#include <vector>

#include "kami_internal.hpp"

namespace kami {

struct SplitArgs {
    const double* A;
    const double* B;
    double* C;
    int64_t sa, sb, sc, lda, ldb, ldc;
    double alpha, beta;
    // persistent form only: K chunks per problem and the grid of BM x BN tiles of C per
    // problem (1 x 1 for a block problem)
    int32_t nch, tiles_m, tiles_n;
    int32_t iters;  // resident form: repetitions of the product
};

typedef hipError_t (*split_launch_fn)(const SplitArgs&, int64_t, hipStream_t);

struct SplitInstance {
    int prec, m, n, k, slices, pm, pn, threads, lds;
    split_launch_fn launch;
    int pipe;      // 1: one workgroup per problem, exponents over full K; 2: persistent
    int resident;  // 1: resident-repeat form, used for iters > 1 only; 2: any iters
};

// The int8-digit instances (split_i8.hip).
void add_split_i8_instances(std::vector<SplitInstance>& v);

}  // namespace kami
