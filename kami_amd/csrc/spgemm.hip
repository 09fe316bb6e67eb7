// Block SpGEMM (placeholder until the symbolic/numeric kernels land).
#include "kami_internal.hpp"

extern "C" {
kami_status kami_bsr_spgemm_workspace_size(int32_t mb, int32_t nb, size_t* bytes) {
    (void)mb; (void)nb; (void)bytes;
    return KAMI_ENOTSUP;
}
kami_status kami_bsr_spgemm_symbolic(const kami_bsr* A, const kami_bsr* B, int32_t* rowptrC,
                                     int64_t* nnzbC, void* ws, size_t ws_bytes, void* stream) {
    (void)A; (void)B; (void)rowptrC; (void)nnzbC; (void)ws; (void)ws_bytes; (void)stream;
    return KAMI_ENOTSUP;
}
kami_status kami_bsr_spgemm_fill(const kami_bsr* A, const kami_bsr* B, const int32_t* rowptrC,
                                 int32_t* colidxC, void* ws, size_t ws_bytes, void* stream) {
    (void)A; (void)B; (void)rowptrC; (void)colidxC; (void)ws; (void)ws_bytes; (void)stream;
    return KAMI_ENOTSUP;
}
kami_status kami_bsr_spgemm_numeric(const kami_bsr* A, const kami_bsr* B, const int32_t* rowptrC,
                                    const int32_t* colidxC, int64_t nnzbC, void* valC,
                                    int32_t precision, double alpha, void* stream) {
    (void)A; (void)B; (void)rowptrC; (void)colidxC; (void)nnzbC; (void)valC; (void)precision;
    (void)alpha; (void)stream;
    return KAMI_ENOTSUP;
}
}
