// FP64 block-GEMM instances (native FP64 MFMA 16x16x4).
// Wave counts follow the reference's FP64 sweep tables
// (scripts/fp64_block_square_GH200.sh:11-25): 1D uses P = M/16, 2D and 3D use P = 1
// up to 32 and P = 2 above; the P = 4 2D variants are extra candidates for the tuner.
#include "block_gemm_registry.hpp"

namespace kami {

void register_f64_instances(std::vector<BlockInstance>& v) {
    using A = AtomF64;
    constexpr int F = 0;  // KAMI_PREC_F64
    // 16^3: a single wave; the three partitions coincide
    add_instance<A, F, 16, 16, 16, 1, 1, 1>(v, 1, 1);
    add_instance<A, F, 16, 16, 16, 1, 1, 1>(v, 2, 1);
    add_instance<A, F, 16, 16, 16, 1, 1, 1>(v, 3, 1);
    // 32^3
    add_instance<A, F, 32, 32, 32, 2, 1, 1>(v, 1, 2);
    add_instance<A, F, 32, 32, 32, 1, 1, 1>(v, 1, 1);
    add_instance<A, F, 32, 32, 32, 1, 1, 1>(v, 2, 1);
    add_instance<A, F, 32, 32, 32, 2, 2, 1>(v, 2, 2);
    add_instance<A, F, 32, 32, 32, 1, 1, 1>(v, 3, 1);
    add_instance<A, F, 32, 32, 32, 2, 2, 2>(v, 3, 2);
    // 64^3
    add_instance<A, F, 64, 64, 64, 4, 1, 1>(v, 1, 4);
    add_instance<A, F, 64, 64, 64, 2, 1, 1>(v, 1, 2);
    add_instance<A, F, 64, 64, 64, 2, 2, 1>(v, 2, 2);
    add_instance<A, F, 64, 64, 64, 4, 4, 1>(v, 2, 4);
    add_instance<A, F, 64, 64, 64, 2, 2, 2>(v, 3, 2);
    // 128^3
    add_instance<A, F, 128, 128, 128, 8, 1, 1>(v, 1, 8);
    add_instance<A, F, 128, 128, 128, 4, 1, 1>(v, 1, 4);
    add_resident_instance<A, F, 128, 128, 128, 8>(v, 8);
    add_instance<A, F, 128, 128, 128, 2, 2, 1>(v, 2, 2);
    add_instance<A, F, 128, 128, 128, 4, 4, 1>(v, 2, 4);
    add_instance<A, F, 128, 128, 128, 2, 2, 2>(v, 3, 2);
    // K beyond one register/LDS stage: K = 512 folded into double-buffered LDS stages
    // (the reference's K fold, src/batched_gemm/KAMI/1d/batched_gemm_1d_double_mma_128.cu:100-205)
    add_instance<A, F, 64, 64, 512, 2, 2, 1>(v, 2, 2);
    add_instance<A, F, 128, 128, 512, 8, 1, 1>(v, 1, 8);
    add_instance<A, F, 128, 128, 512, 4, 4, 1>(v, 2, 4);
    // persistent 2D forms (pipeline 2)
    add_pipe_instance<A, F, 32, 32, 32, 2, 2, 4>(v, 2);
    add_pipe_instance<A, F, 64, 64, 64, 2, 2, 4>(v, 2);
    add_pipe_instance<A, F, 64, 64, 64, 4, 4, 4>(v, 4);
    add_pipe_instance<A, F, 128, 128, 128, 2, 2, 4>(v, 2);
    add_pipe_instance<A, F, 128, 128, 128, 4, 4, 2>(v, 4);
}

}  // namespace kami
