/*
 * kami_oracle.c -- CPU oracle for the block-GEMM / batched GEMM / block-SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the host-side algorithms
 * of the reference (paths relative to the KAMI source tree) so that the HIP library
 * can be checked against them.  Only tests/, __graft_entry__.smoke() and the CPU
 * baseline leg of bench.py load it.  Nothing in the library calls it.
 *
 * Every function names the reference lines it follows.  Checks that pin each one
 * independently of this file (closed forms, brute force, textbook identities) live in
 * tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "kami_oracle.h"

/* ------------------------------------------------------------------ scalar formats */

/* IEEE binary16 -> double, from the format definition (1 sign, 5 exponent, 10 fraction
 * bits, bias 15; exponent 0 = subnormal, 31 = inf/nan). */
double oracle_half_to_double(uint16_t h) {
    const int s = (h >> 15) & 1, e = (h >> 10) & 0x1f, f = h & 0x3ff;
    double v;
    if (e == 0) v = ldexp((double)f, -24);
    else if (e == 31) v = f ? NAN : INFINITY;
    else v = ldexp((double)(f | 0x400), e - 25);
    return s ? -v : v;
}

/* double -> binary16 with round-to-nearest-even, overflow to inf. */
uint16_t oracle_double_to_half(double x) {
    uint16_t sign = 0;
    if (signbit(x)) { sign = 0x8000; x = -x; }
    if (isnan(x)) return 0x7e00;
    if (x >= 65520.0) return sign | 0x7c00; /* rounds to inf */
    if (x < ldexp(1.0, -25)) {
        /* below half the smallest subnormal: 0, except exact ties handled below */
        return sign;
    }
    int e;
    frexp(x, &e); /* x = m * 2^e, m in [0.5, 1) */
    int exp_unb = e - 1; /* x in [2^exp_unb, 2^(exp_unb+1)) */
    if (exp_unb < -14) exp_unb = -14; /* subnormal range shares the quantum 2^-24 */
    double q = ldexp(1.0, exp_unb - 10);
    double m = nearbyint(x / q); /* default rounding mode is nearest-even */
    double v = m * q;
    /* re-normalise if rounding carried into the next binade */
    int ee;
    frexp(v, &ee);
    int ue = ee - 1;
    if (v == 0) return sign;
    if (ue > 15) return sign | 0x7c00;
    if (ue < -14) { /* subnormal */
        return sign | (uint16_t)(v / ldexp(1.0, -24));
    }
    uint16_t frac = (uint16_t)(v / ldexp(1.0, ue - 10)) - 0x400;
    return sign | (uint16_t)((ue + 15) << 10) | frac;
}

double oracle_bf16_to_double(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return (double)f;
}

uint16_t oracle_double_to_bf16(double x) {
    float f = (float)x; /* double -> float (RNE), then float -> bf16 (RNE) */
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

/* ------------------------------------------------------------------ dense GEMM */

static inline double load_b(const double* B, int64_t ldb, int b_kn, int k, int j) {
    return b_kn ? B[(int64_t)k * ldb + j] : B[(int64_t)j * ldb + k];
}

/* C(i,j) = sum_k A(i,k) B(k,j): the reference's validation triple loop, loop order
 * i, j, k with B stored N x K (src/block_gemm/KAMI/1d/fp64/block_gemm_1d_double_mma.cu:
 * 294-304); b_kn selects the K x N storage used by the AMD port
 * (src/block_gemm/amd/2d/block_gemm_2d_half_wmma.cu:266). */
void oracle_gemm_f64(int m, int n, int k, const double* A, int64_t lda, const double* B,
                     int64_t ldb, int b_kn, double* C, int64_t ldc) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int kk = 0; kk < k; kk++) s += A[(int64_t)i * lda + kk] * load_b(B, ldb, b_kn, kk, j);
            C[(int64_t)i * ldc + j] = s;
        }
}

/* One FP64 triple loop per problem with OpenMP over the batch
 * (src/batched_gemm/KAMI/1d/batched_gemm_1d_double_mma.cu:304-318). */
void oracle_batched_gemm_f64(int64_t count, int m, int n, int k, const double* A, int64_t sa,
                             const double* B, int64_t sb, double* C, int64_t sc, int b_kn,
                             int nthreads) {
    const int64_t ldb = b_kn ? n : k;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t p = 0; p < count; p++)
        oracle_gemm_f64(m, n, k, A + p * sa, k, B + p * sb, ldb, b_kn, C + p * sc, n);
    (void)nthreads;
}

/* The same loop over problems of their own sizes (the uneven batched workload): problem p
 * is mnk[3p] x mnk[3p+1] x mnk[3p+2], its A, B (N x K) and C packed one after another in
 * problem order.  Dynamic scheduling balances the uneven problems over the threads. */
void oracle_batched_uneven_f64(int64_t count, const int32_t* mnk, const double* A,
                               const double* B, double* C, int nthreads) {
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * 3 * (size_t)(count + 1));
    if (!off) return;
    off[0] = off[1] = off[2] = 0;
    for (int64_t p = 0; p < count; p++) {
        const int64_t m = mnk[3 * p], n = mnk[3 * p + 1], k = mnk[3 * p + 2];
        off[3 * p + 3] = off[3 * p] + m * k;
        off[3 * p + 4] = off[3 * p + 1] + n * k;
        off[3 * p + 5] = off[3 * p + 2] + m * n;
    }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads)
#endif
    for (int64_t p = 0; p < count; p++) {
        const int m = mnk[3 * p], n = mnk[3 * p + 1], k = mnk[3 * p + 2];
        oracle_gemm_f64(m, n, k, A + off[3 * p], k, B + off[3 * p + 1], k, 0, C + off[3 * p + 2], n);
    }
    free(off);
    (void)nthreads;
}

/* FP16 / BF16 inputs widened to double and accumulated in double, the reference's
 * FP16 validation loop (src/block_gemm/KAMI/1d/fp16/block_gemm_1d_half_mma.cu:
 * 285-296) with its (double) casts. */
static void gemm_u16(int m, int n, int k, const uint16_t* A, int64_t lda, const uint16_t* B,
                     int64_t ldb, int b_kn, double* C, int64_t ldc, double (*cv)(uint16_t)) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int kk = 0; kk < k; kk++) {
                uint16_t b = b_kn ? B[(int64_t)kk * ldb + j] : B[(int64_t)j * ldb + kk];
                s += cv(A[(int64_t)i * lda + kk]) * cv(b);
            }
            C[(int64_t)i * ldc + j] = s;
        }
}

void oracle_gemm_f16(int m, int n, int k, const uint16_t* A, int64_t lda, const uint16_t* B,
                     int64_t ldb, int b_kn, double* C, int64_t ldc) {
    gemm_u16(m, n, k, A, lda, B, ldb, b_kn, C, ldc, oracle_half_to_double);
}

void oracle_gemm_bf16(int m, int n, int k, const uint16_t* A, int64_t lda, const uint16_t* B,
                      int64_t ldb, int b_kn, double* C, int64_t ldc) {
    gemm_u16(m, n, k, A, lda, B, ldb, b_kn, C, ldc, oracle_bf16_to_double);
}

void oracle_batched_gemm_u16(int64_t count, int m, int n, int k, const uint16_t* A, int64_t sa,
                             const uint16_t* B, int64_t sb, double* C, int64_t sc, int b_kn,
                             int is_bf16, int nthreads) {
    const int64_t ldb = b_kn ? n : k;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t p = 0; p < count; p++)
        gemm_u16(m, n, k, A + p * sa, k, B + p * sb, ldb, b_kn, C + p * sc, n,
                 is_bf16 ? oracle_bf16_to_double : oracle_half_to_double);
    (void)nthreads;
}

/* Compensated dot product Dot2 (Ogita, Rump, Oishi, "Accurate sum and dot product",
 * SIAM J. Sci. Comput. 26(6), 2005, Algorithm 5.3): TwoProduct via fma, TwoSum
 * (Knuth) for the running sum, errors accumulated and added at the end.  The result
 * is as accurate as if computed in twice the working precision, so it serves as the
 * exact reference for the 1e-13 error checks.  Also returns sum |a_k b_k| in
 * absC (may be NULL) for component-wise relative errors. */
void oracle_gemm_dot2_f64(int m, int n, int k, const double* A, int64_t lda, const double* B,
                          int64_t ldb, int b_kn, double* C, int64_t ldc, double* absC) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            double p = 0, s = 0, absum = 0;
            for (int kk = 0; kk < k; kk++) {
                const double a = A[(int64_t)i * lda + kk], b = load_b(B, ldb, b_kn, kk, j);
                const double h = a * b;
                const double r = fma(a, b, -h);      /* TwoProduct error */
                const double t = p + h;              /* TwoSum */
                const double z = t - p;
                const double q = (p - (t - z)) + (h - z);
                p = t;
                s += q + r;
                absum += fabs(h);
            }
            C[(int64_t)i * ldc + j] = p + s;
            if (absC) absC[(int64_t)i * ldc + j] = absum;
        }
}

void oracle_batched_gemm_dot2_f64(int64_t count, int m, int n, int k, const double* A,
                                  int64_t sa, const double* B, int64_t sb, double* C,
                                  int64_t sc, double* absC, int b_kn, int nthreads) {
    const int64_t ldb = b_kn ? n : k;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t p = 0; p < count; p++)
        oracle_gemm_dot2_f64(m, n, k, A + p * sa, k, B + p * sb, ldb, b_kn, C + p * sc, n,
                             absC ? absC + p * sc : NULL);
    (void)nthreads;
}

/* 3D slab semantics: C = sum over d of the P partial slabs stored at C + d*m*n
 * (src/block_gemm/KAMI/3d/fp64/block_gemm_3d_double_mma.cu:324-333). */
void oracle_slab_sum_f64(int m, int n, int p, double* C) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++)
            for (int d = 1; d < p; d++) C[i * n + j] += C[i * n + j + (int64_t)d * m * n];
}

/* ------------------------------------------------------------------ generators */

/* Ramp known-answer input X[i] = i (src/block_gemm/KAMI/1d/fp64/
 * block_gemm_1d_double_mma.cu:224-236). */
void oracle_fill_ramp(double* X, int64_t n) {
    for (int64_t i = 0; i < n; i++) X[i] = (double)i;
}

/* X[i] = rand_r(&seed) % 3 with seed = i, per element (src/batched_gemm/KAMI/1d/
 * batched_gemm_1d_double_mma.cu:231-245).  rand_r is the C library's. */
void oracle_fill_randr_mod3(double* X, int64_t n) {
    for (int64_t i = 0; i < n; i++) {
        unsigned int seed = (unsigned int)i;
        X[i] = (double)(rand_r(&seed) % 3);
    }
}

/* Random BSR matrix (src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:372-397):
 * srand(0); for every block (row-major order) keep it when rand()/RAND_MAX >
 * sparsity, then draw its bs*bs values as rand()%3+1 (stored here as binary16 bits).
 * Returns the number of stored blocks (the reference returns elements via *nnz). */
int oracle_bsr_generate(int rows, int cols, float sparsity, int bs, int32_t* row_ptr,
                        int32_t* col_idx, uint16_t* values) {
    int count = 0;
    srand(0);
    row_ptr[0] = 0;
    for (int i = 0; i < rows; i++) {
        for (int j = 0; j < cols; j++) {
            float r = (float)rand() / (float)RAND_MAX;
            if (r > sparsity) {
                col_idx[count] = j;
                for (int len = 0; len < bs * bs; len++)
                    values[(int64_t)count * bs * bs + len] = oracle_double_to_half((double)(rand() % 3) + 1.0);
                count++;
            }
        }
        row_ptr[i + 1] = count;
    }
    return count;
}

/* ------------------------------------------------------------------ block SpGEMM */

/* In-place exclusive scan over `length` entries with the reference's recurrence
 * (src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:269-283). */
void oracle_exclusive_scan_i32(int32_t* x, int length) {
    if (length == 0 || length == 1) return;
    int32_t old_val = x[0], new_val;
    x[0] = 0;
    for (int i = 1; i < length; i++) {
        new_val = x[i];
        x[i] = old_val + x[i - 1];
        old_val = new_val;
    }
}

/* Symbolic phase: per block row a dense flag array over B's columns, marked through
 * A's columns, counted, then exclusive-scanned into RowPtrC[0..mb]
 * (src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:289-317). Returns nnzb(C). */
int64_t oracle_spgemm_symbolic(int mb, int nb, const int32_t* rpA, const int32_t* ciA,
                               const int32_t* rpB, const int32_t* ciB, int32_t* rpC) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int iid = 0; iid < mb; iid++) {
        char* flag = (char*)calloc((size_t)nb, 1);
        for (int i = rpA[iid]; i < rpA[iid + 1]; i++) {
            const int col = ciA[i];
            for (int l = rpB[col]; l < rpB[col + 1]; l++) flag[ciB[l]] = 1;
        }
        int nnzr = 0;
        for (int c = 0; c < nb; c++) nnzr += flag[c] == 1;
        rpC[iid] = nnzr;
        free(flag);
    }
    rpC[mb] = 0;
    oracle_exclusive_scan_i32(rpC, mb + 1);
    return rpC[mb];
}

/* Column fill in ascending order (src/sparse/spgemm/2d/block_spgemm_2d_half_mma.cu:
 * 88-116). */
void oracle_spgemm_fill(int mb, int nb, const int32_t* rpA, const int32_t* ciA,
                        const int32_t* rpB, const int32_t* ciB, const int32_t* rpC,
                        int32_t* ciC) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int iid = 0; iid < mb; iid++) {
        char* flag = (char*)calloc((size_t)nb, 1);
        for (int i = rpA[iid]; i < rpA[iid + 1]; i++) {
            const int col = ciA[i];
            for (int l = rpB[col]; l < rpB[col + 1]; l++) flag[ciB[l]] = 1;
        }
        int nnzr = rpC[iid];
        for (int c = 0; c < nb; c++)
            if (flag[c] == 1) ciC[nnzr++] = c;
        free(flag);
    }
}

/* Numeric phase with a dense per-row accumulator (SPA) over B's column blocks,
 * values widened to double (src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:321-368;
 * the reference stores the result as half, this oracle keeps double).  `is_bf16`
 * selects the input format; values are bs x bs row-major blocks. */
void oracle_spgemm_numeric_u16(int mb, int nb, int bs, const int32_t* rpA, const int32_t* ciA,
                               const uint16_t* vA, const int32_t* rpB, const int32_t* ciB,
                               const uint16_t* vB, const int32_t* rpC, double* vC, int is_bf16) {
    double (*cv)(uint16_t) = is_bf16 ? oracle_bf16_to_double : oracle_half_to_double;
    const int64_t b2 = (int64_t)bs * bs;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int iid = 0; iid < mb; iid++) {
        char* flag = (char*)calloc((size_t)nb, 1);
        double* row = (double*)calloc((size_t)nb * b2, sizeof(double));
        for (int i = rpA[iid]; i < rpA[iid + 1]; i++) {
            const int col = ciA[i];
            const uint16_t* a = vA + (int64_t)i * b2;
            for (int l = rpB[col]; l < rpB[col + 1]; l++) {
                const int key = ciB[l];
                const uint16_t* b = vB + (int64_t)l * b2;
                flag[key] = 1;
                double* c = row + (int64_t)key * b2;
                for (int ii = 0; ii < bs; ii++)
                    for (int jj = 0; jj < bs; jj++) {
                        double s = 0;
                        for (int kk = 0; kk < bs; kk++) s += cv(a[ii * bs + kk]) * cv(b[kk * bs + jj]);
                        c[ii * bs + jj] += s;
                    }
            }
        }
        int nnzr = rpC[iid];
        for (int c = 0; c < nb; c++)
            if (flag[c] == 1) {
                memcpy(vC + (int64_t)nnzr * b2, row + (int64_t)c * b2, (size_t)b2 * sizeof(double));
                nnzr++;
            }
        free(flag);
        free(row);
    }
}

/* FP64-valued variant of the numeric phase (same algorithm). */
void oracle_spgemm_numeric_f64(int mb, int nb, int bs, const int32_t* rpA, const int32_t* ciA,
                               const double* vA, const int32_t* rpB, const int32_t* ciB,
                               const double* vB, const int32_t* rpC, double* vC) {
    const int64_t b2 = (int64_t)bs * bs;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int iid = 0; iid < mb; iid++) {
        char* flag = (char*)calloc((size_t)nb, 1);
        double* row = (double*)calloc((size_t)nb * b2, sizeof(double));
        for (int i = rpA[iid]; i < rpA[iid + 1]; i++) {
            const int col = ciA[i];
            const double* a = vA + (int64_t)i * b2;
            for (int l = rpB[col]; l < rpB[col + 1]; l++) {
                const int key = ciB[l];
                const double* b = vB + (int64_t)l * b2;
                flag[key] = 1;
                double* c = row + (int64_t)key * b2;
                for (int ii = 0; ii < bs; ii++)
                    for (int jj = 0; jj < bs; jj++) {
                        double s = 0;
                        for (int kk = 0; kk < bs; kk++) s += a[ii * bs + kk] * b[kk * bs + jj];
                        c[ii * bs + jj] += s;
                    }
            }
        }
        int nnzr = rpC[iid];
        for (int c = 0; c < nb; c++)
            if (flag[c] == 1) {
                memcpy(vC + (int64_t)nnzr * b2, row + (int64_t)c * b2, (size_t)b2 * sizeof(double));
                nnzr++;
            }
        free(flag);
        free(row);
    }
}

/* Number of block products, the reference's FLOP unit for SpGEMM
 * (src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:442-447). */
int64_t oracle_spgemm_nnzcub(int nnzbA, const int32_t* ciA, const int32_t* rpB) {
    int64_t s = 0;
    for (int i = 0; i < nnzbA; i++) s += rpB[ciA[i] + 1] - rpB[ciA[i]];
    return s;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Block SpMM check loop (src/sparse/spmm/2d/block_spmm_2d_half_mma.cu:378-403) on plain BSR
 * (the reference walks its per-rank re-tiling of the same blocks): for every stored block
 * (i, kb) and every r, c of the block and column j of C,
 *   C[(i bs + r) ldc + j] += A(i, kb)[r][c] * B[j ldb + kb bs + c]       (b_kn = 0, B n x K)
 *   C[(i bs + r) ldc + j] += A(i, kb)[r][c] * B[(kb bs + c) ldb + j]     (b_kn = 1, B K x n)
 * in double.  C is accumulated into (the caller zeroes it). */
void oracle_bsr_spmm_f64(int mb, int nb, int bs, const int32_t* rp, const int32_t* ci,
                         const double* vals, const double* B, int n, int64_t ldb, int b_kn,
                         double* C, int64_t ldc) {
    (void)nb;
    const int64_t b2 = (int64_t)bs * bs;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int i = 0; i < mb; i++)
        for (int q = rp[i]; q < rp[i + 1]; q++) {
            const double* a = vals + (int64_t)q * b2;
            const int64_t kb = ci[q];
            for (int r = 0; r < bs; r++)
                for (int c = 0; c < bs; c++) {
                    const double x = a[r * bs + c];
                    double* crow = C + ((int64_t)i * bs + r) * ldc;
                    for (int j = 0; j < n; j++) {
                        const double y = b_kn ? B[(kb * bs + c) * ldb + j] : B[j * ldb + kb * bs + c];
                        crow[j] += x * y;
                    }
                }
        }
}

/* The SpMM program's inputs (src/sparse/spmm/2d/block_spmm_2d_half_mma.cu:237-262 and
 * :315-320): srand(0); a block is kept when rand()/RAND_MAX > sparsity and its values are
 * rand() % 2 + 1; then the n x K dense B takes rand() % 2 + 1 continuing the same
 * sequence.  Returns the number of stored blocks. */
int oracle_spmm_generate(int rows, int cols, float sparsity, int bs, int32_t* rp, int32_t* ci,
                         double* vals, int n, double* B) {
    int count = 0;
    srand(0);
    rp[0] = 0;
    for (int i = 0; i < rows; i++) {
        for (int j = 0; j < cols; j++) {
            float r = (float)rand() / (float)RAND_MAX;
            if (r > sparsity) {
                ci[count] = j;
                for (int len = 0; len < bs * bs; len++)
                    vals[(int64_t)count * bs * bs + len] = (double)(rand() % 2) + 1.0;
                count++;
            }
        }
        rp[i + 1] = count;
    }
    const int64_t nb_elems = (int64_t)n * cols * bs;
    for (int64_t i = 0; i < nb_elems; i++) B[i] = (double)(rand() % 2) + 1.0;
    return count;
}
