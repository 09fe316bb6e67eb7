/* kami_oracle.h -- CPU oracle (test infrastructure only; see kami_oracle.c). */
#ifndef KAMI_ORACLE_H
#define KAMI_ORACLE_H
#include <stdint.h>

double oracle_half_to_double(uint16_t h);
uint16_t oracle_double_to_half(double x);
double oracle_bf16_to_double(uint16_t b);
uint16_t oracle_double_to_bf16(double x);

void oracle_gemm_f64(int m, int n, int k, const double* A, int64_t lda, const double* B,
                     int64_t ldb, int b_kn, double* C, int64_t ldc);
void oracle_batched_gemm_f64(int64_t count, int m, int n, int k, const double* A, int64_t sa,
                             const double* B, int64_t sb, double* C, int64_t sc, int b_kn,
                             int nthreads);
void oracle_gemm_f16(int m, int n, int k, const uint16_t* A, int64_t lda, const uint16_t* B,
                     int64_t ldb, int b_kn, double* C, int64_t ldc);
void oracle_gemm_bf16(int m, int n, int k, const uint16_t* A, int64_t lda, const uint16_t* B,
                      int64_t ldb, int b_kn, double* C, int64_t ldc);
void oracle_batched_uneven_f64(int64_t count, const int32_t* mnk, const double* A,
                               const double* B, double* C, int nthreads);
void oracle_batched_gemm_u16(int64_t count, int m, int n, int k, const uint16_t* A, int64_t sa,
                             const uint16_t* B, int64_t sb, double* C, int64_t sc, int b_kn,
                             int is_bf16, int nthreads);
void oracle_gemm_dot2_f64(int m, int n, int k, const double* A, int64_t lda, const double* B,
                          int64_t ldb, int b_kn, double* C, int64_t ldc, double* absC);
void oracle_batched_gemm_dot2_f64(int64_t count, int m, int n, int k, const double* A,
                                  int64_t sa, const double* B, int64_t sb, double* C,
                                  int64_t sc, double* absC, int b_kn, int nthreads);
void oracle_slab_sum_f64(int m, int n, int p, double* C);

void oracle_fill_ramp(double* X, int64_t n);
void oracle_fill_randr_mod3(double* X, int64_t n);
int oracle_bsr_generate(int rows, int cols, float sparsity, int bs, int32_t* row_ptr,
                        int32_t* col_idx, uint16_t* values);

void oracle_exclusive_scan_i32(int32_t* x, int length);
int64_t oracle_spgemm_symbolic(int mb, int nb, const int32_t* rpA, const int32_t* ciA,
                               const int32_t* rpB, const int32_t* ciB, int32_t* rpC);
void oracle_spgemm_fill(int mb, int nb, const int32_t* rpA, const int32_t* ciA,
                        const int32_t* rpB, const int32_t* ciB, const int32_t* rpC,
                        int32_t* ciC);
void oracle_spgemm_numeric_u16(int mb, int nb, int bs, const int32_t* rpA, const int32_t* ciA,
                               const uint16_t* vA, const int32_t* rpB, const int32_t* ciB,
                               const uint16_t* vB, const int32_t* rpC, double* vC, int is_bf16);
void oracle_spgemm_numeric_f64(int mb, int nb, int bs, const int32_t* rpA, const int32_t* ciA,
                               const double* vA, const int32_t* rpB, const int32_t* ciB,
                               const double* vB, const int32_t* rpC, double* vC);
int64_t oracle_spgemm_nnzcub(int nnzbA, const int32_t* ciA, const int32_t* rpB);
void oracle_bsr_spmm_f64(int mb, int nb, int bs, const int32_t* rp, const int32_t* ci,
                         const double* vals, const double* B, int n, int64_t ldb, int b_kn,
                         double* C, int64_t ldc);
int oracle_spmm_generate(int rows, int cols, float sparsity, int bs, int32_t* rp, int32_t* ci,
                         double* vals, int n, double* B);
int oracle_max_threads(void);
#endif
