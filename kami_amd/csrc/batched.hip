// Batched GEMM with a persistent kernel (placeholder until the queue kernel lands).
#include "kami_internal.hpp"

extern "C" {
kami_status kami_batched_plan_create(const kami_problem* probs, int64_t count, int32_t precision,
                                     const kami_batched_opts* opts, void* stream,
                                     kami_batched_plan** out) {
    (void)probs; (void)count; (void)precision; (void)opts; (void)stream; (void)out;
    return KAMI_ENOTSUP;
}
kami_status kami_batched_execute(kami_batched_plan* plan, void* stream) {
    (void)plan; (void)stream;
    return KAMI_ENOTSUP;
}
kami_status kami_batched_plan_info(const kami_batched_plan* plan, kami_plan_info* out) {
    (void)plan; (void)out;
    return KAMI_ENOTSUP;
}
kami_status kami_batched_plan_destroy(kami_batched_plan* plan) {
    (void)plan;
    return KAMI_OK;
}
}
