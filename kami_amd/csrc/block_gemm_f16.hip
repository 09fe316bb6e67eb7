// Block-GEMM instances: FP16 inputs, FP32 accumulation (16x16x32, 16x16x16 and 32x32x16 f16
// MFMA).
// Shapes are the reference's FP16 sweep (scripts/fp16_block_square_GH200.sh:11-25,
// scripts/fp16_block_lowrank_GH200.sh:20-27): squares {16, 32, 64, 128, 192, 256} and
// low-rank M = N in {32, 64, 128, 192, 256} with K in {16, 32}.  The 3D partition at
// M = N >= 192 runs in column passes (add_pass_instance): a full-width per-wave FP32
// accumulator tile at two depths exceeds the register file at every legal workgroup
// size; K = 16 splits over two depths with a K step of 8 on the 16x16x16 instruction.
#include "block_gemm_registry.hpp"

namespace kami {

void register_f16_instances(std::vector<BlockInstance>& v) {
    using A = AtomF16x32;
    using H = AtomF16x16;  // K step 16, for K slices that are not multiples of 32
    using W = AtomF16M32;  // 32x32x16: full rate at K = 16, wave tiles multiples of 32
    using H8 = AtomF16x16K8;  // K step 8: the 3D split of K = 16
    constexpr int F = 1;
    // 32^3
    add_instance<A, F, 32, 32, 32, 1, 1, 1>(v, 1, 1);
    add_instance<A, F, 32, 32, 32, 2, 1, 1>(v, 1, 2);
    add_instance<A, F, 32, 32, 32, 1, 1, 1>(v, 2, 1);
    add_instance<A, F, 32, 32, 32, 2, 2, 1>(v, 2, 2);
    add_instance<A, F, 32, 32, 32, 1, 1, 1>(v, 3, 1);
    add_instance<H, F, 32, 32, 32, 2, 2, 2>(v, 3, 2);
    // 64^3
    add_instance<A, F, 64, 64, 64, 1, 1, 1>(v, 1, 1);
    add_instance<A, F, 64, 64, 64, 2, 1, 1>(v, 1, 2);
    add_instance<A, F, 64, 64, 64, 4, 1, 1>(v, 1, 4);
    add_instance<A, F, 64, 64, 64, 1, 1, 1>(v, 2, 1);
    add_instance<A, F, 64, 64, 64, 2, 2, 1>(v, 2, 2);
    add_instance<A, F, 64, 64, 64, 4, 4, 1>(v, 2, 4);
    add_instance<A, F, 64, 64, 64, 1, 1, 1>(v, 3, 1);
    add_instance<A, F, 64, 64, 64, 2, 2, 2>(v, 3, 2);
    // 128^3
    add_instance<A, F, 128, 128, 128, 4, 1, 1>(v, 1, 4);
    add_instance<A, F, 128, 128, 128, 8, 1, 1>(v, 1, 8);
    add_instance<A, F, 128, 128, 128, 2, 2, 1>(v, 2, 2);
    add_instance<A, F, 128, 128, 128, 4, 4, 1>(v, 2, 4);
    add_instance<A, F, 128, 128, 128, 2, 2, 2>(v, 3, 2);
    // 256^3
    add_instance<A, F, 256, 256, 256, 8, 1, 1>(v, 1, 8);
    add_instance<A, F, 256, 256, 256, 16, 1, 1>(v, 1, 16);
    add_resident_instance<A, F, 256, 256, 256, 8>(v, 8);
    add_resident_instance<A, F, 256, 256, 256, 16>(v, 16);
    add_instance<A, F, 256, 256, 256, 4, 4, 1>(v, 2, 4);
    // low rank 256 x 256 x 16 and x 32
    add_instance<W, F, 256, 256, 16, 8, 1, 1>(v, 1, 8);
    add_instance<H, F, 256, 256, 16, 16, 1, 1>(v, 1, 16);
    add_instance<W, F, 256, 256, 16, 4, 4, 1>(v, 2, 4);
    add_instance<A, F, 256, 256, 32, 8, 1, 1>(v, 1, 8);
    add_instance<A, F, 256, 256, 32, 16, 1, 1>(v, 1, 16);
    add_instance<A, F, 256, 256, 32, 4, 4, 1>(v, 2, 4);
    // persistent 2D forms (pipeline 2)
    // 16^3 (one wave) and 192^3 (the largest size of the reference's FP16 square sweep,
    // scripts/fp16_block_square_GH200.sh:13-26; 2D only: 1D and 3D tiles at 192 spill)
    add_instance<H, F, 16, 16, 16, 1, 1, 1>(v, 1, 1);
    add_instance<H, F, 16, 16, 16, 1, 1, 1>(v, 2, 1);
    add_instance<H, F, 16, 16, 16, 1, 1, 1>(v, 3, 1);
    add_instance<A, F, 192, 192, 192, 4, 4, 1>(v, 2, 4);
    add_resident_instance<A, F, 192, 192, 192, 4, 4>(v, 4);
    // 1D, 12 waves of 16 rows: all of K fits one stage, so this form is resident as is
    add_instance<A, F, 192, 192, 192, 12, 1, 1, 2>(v, 1, 12);
    // low-rank M = N in {32, 64, 128, 192}, K in {16, 32}
    // (scripts/fp16_block_lowrank_GH200.sh:20-27)
    add_instance<H, F, 32, 32, 16, 1, 1, 1>(v, 1, 1);
    add_instance<H, F, 32, 32, 16, 2, 1, 1>(v, 1, 2);
    add_instance<H, F, 32, 32, 16, 1, 1, 1>(v, 2, 1);
    add_instance<H, F, 32, 32, 16, 2, 2, 1>(v, 2, 2);
    add_instance<H, F, 64, 64, 16, 1, 1, 1>(v, 1, 1);
    add_instance<H, F, 64, 64, 16, 2, 1, 1>(v, 1, 2);
    add_instance<H, F, 64, 64, 16, 4, 1, 1>(v, 1, 4);
    add_instance<H, F, 64, 64, 16, 2, 2, 1>(v, 2, 2);
    add_instance<A, F, 64, 64, 32, 1, 1, 1>(v, 1, 1);
    add_instance<A, F, 64, 64, 32, 2, 1, 1>(v, 1, 2);
    add_instance<A, F, 64, 64, 32, 4, 1, 1>(v, 1, 4);
    add_instance<A, F, 64, 64, 32, 2, 2, 1>(v, 2, 2);
    add_instance<H, F, 128, 128, 16, 2, 1, 1>(v, 1, 2);
    add_instance<W, F, 128, 128, 16, 4, 1, 1>(v, 1, 4);
    add_instance<H, F, 128, 128, 16, 8, 1, 1>(v, 1, 8);
    add_instance<W, F, 128, 128, 16, 2, 2, 1>(v, 2, 2);
    add_instance<A, F, 128, 128, 32, 2, 1, 1>(v, 1, 2);
    add_instance<A, F, 128, 128, 32, 4, 1, 1>(v, 1, 4);
    add_instance<A, F, 128, 128, 32, 8, 1, 1>(v, 1, 8);
    add_instance<A, F, 128, 128, 32, 2, 2, 1>(v, 2, 2);
    add_instance<H, F, 192, 192, 16, 4, 1, 1>(v, 1, 4);
    add_instance<H, F, 192, 192, 16, 12, 1, 1>(v, 1, 12);
    add_instance<H, F, 192, 192, 16, 4, 4, 1>(v, 2, 4);
    add_instance<A, F, 192, 192, 32, 4, 1, 1>(v, 1, 4);
    add_instance<A, F, 192, 192, 32, 12, 1, 1>(v, 1, 12);
    add_instance<A, F, 192, 192, 32, 4, 4, 1>(v, 2, 4);

    // 208^3 and 224^3 (the reference's FP16 square sweep reaches 224,
    // logs/block_gemm/square/H200/fp16_block_square_1d_H200.csv:13-14): 1D, 13 and 14
    // waves of 16 rows; B's whole K (90-107 KB) exceeds the stage budget, so K is staged
    // 16 (208) or 32 (224) wide through double-buffered LDS
    add_instance<H, F, 208, 208, 208, 13, 1, 1, 2>(v, 1, 13);  // K step 16: 208 = 6.5 x 32
    add_instance<A, F, 224, 224, 224, 14, 1, 1, 2>(v, 1, 14);
    // K = 512 folded into LDS stages (src/batched_gemm/KAMI/1d/batched_gemm_1d_double_mma_128.cu:100-205)
    add_instance<A, F, 128, 128, 512, 8, 1, 1>(v, 1, 8);
    add_instance<A, F, 128, 128, 512, 4, 4, 1>(v, 2, 4);
    add_instance<A, F, 256, 256, 512, 16, 1, 1>(v, 1, 16);

    // 3D at 192^3, 256^3 and 256^2 x {16, 32} (2 x 2 x 2 waves) in column passes: the
    // full-width wave tiles (96 x 96 and 128 x 128 FP32 per wave at two depths) exceed
    // the register file, tiles of BN / SP columns do not
    add_pass_instance<A, F, 192, 192, 192, 2, 2, 2, 2>(v, 2);
    add_pass_instance<A, F, 256, 256, 256, 2, 2, 2, 4>(v, 2);
    add_pass_instance<H, F, 256, 256, 32, 2, 2, 2, 2>(v, 2);
    add_pass_instance<H8, F, 256, 256, 16, 2, 2, 2, 2>(v, 2);

    add_pipe_instance<A, F, 64, 64, 64, 2, 2, 4>(v, 2);
    add_pipe_instance<A, F, 64, 64, 64, 4, 4, 4>(v, 4);
    add_pipe_instance<A, F, 128, 128, 128, 2, 2, 4>(v, 2);
    add_pipe_instance<A, F, 128, 128, 128, 4, 4, 4>(v, 4);
    add_pipe_instance<H, F, 256, 256, 16, 4, 4, 4>(v, 4);
}

}  // namespace kami
