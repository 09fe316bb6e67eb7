// Internal declarations shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/kami.h"
#include "block_gemm_kernel.hpp"
#include "block_gemm_registry.hpp"

namespace kami {

extern thread_local int32_t g_last_hip;
kami_status hip_fail(hipError_t e);
kami_status device_check();
const std::vector<BlockInstance>& registry();
kami_status validate_desc(const kami_gemm_desc* d);
const BlockInstance* resolve_block(const kami_gemm_desc* d, kami_status* st);
void set_tuned_choice(int prec, int m, int n, int k, int part, int ranks, int pipe);
// Block-GEMM choices recorded by the autotuner (or loaded from a tuning file).
struct TunedBlock {
    int prec, m, n, k, part, ranks, pipe;
};
std::vector<TunedBlock> tuned_block_choices();
void clear_tuned_block_choices();
// Batched tile variant recorded per (precision, size bucket); -1 when none.
constexpr int kBatchedBuckets = 125;
constexpr int kBatchedVariants = 6;
int batched_tuned_variant(int prec, int bucket);
void set_batched_tuned_variant(int prec, int bucket, int variant);
void clear_batched_tuned_variants();
int elt_in(int prec);
int elt_out(int prec);

// Largest byte span one operand (or one batch of operands) may describe: 2^48 bytes,
// far beyond any device's memory.  Validators reject larger shapes before multiplying,
// so every offset computed from a validated shape stays inside int64.
constexpr int64_t kMaxSpanBytes = (int64_t)1 << 48;
// rows x ld elements of elt bytes within kMaxSpanBytes; rows >= 1, ld >= 0, elt >= 1
inline bool span_fits(int64_t rows, int64_t ld, int64_t elt) {
    return rows >= 1 && ld >= 0 && ld <= kMaxSpanBytes / elt / rows;
}

inline bool is_split(int prec) {
    return prec == KAMI_PREC_F64_SPLIT_F16 || prec == KAMI_PREC_F64_SPLIT_BF16 ||
           prec == KAMI_PREC_F64_SPLIT_I8;
}
kami_status split_query(const kami_gemm_desc* d, kami_gemm_config* out);
kami_status split_launch(const kami_gemm_desc* d, const BlockArgs& args, int64_t count,
                         hipStream_t stream);
kami_status split_gemm_tiled(const kami_gemm_desc* d, const BlockArgs& args, int64_t count,
                             hipStream_t stream);

}  // namespace kami
