// Split-precision FP64 block GEMM (placeholder until the slicing kernels land).
#include "kami_internal.hpp"

namespace kami {

kami_status split_query(const kami_gemm_desc* d, kami_gemm_config* out) {
    (void)d;
    (void)out;
    return KAMI_ENOTSUP;
}

kami_status split_launch(const kami_gemm_desc* d, const BlockArgs& args, int64_t count,
                         hipStream_t stream) {
    (void)d;
    (void)args;
    (void)count;
    (void)stream;
    return KAMI_ENOTSUP;
}

}  // namespace kami
