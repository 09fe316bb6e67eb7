/*
 * kami.h -- C ABI of the MI355X (gfx950) block-level GEMM library.
 *
 * Three entry-point families, all device-pointer based and caller-owned:
 *   block GEMM      kami_block_gemm            one problem per workgroup, fixed shape
 *   batched GEMM    kami_batched_plan_*        uneven shapes, persistent kernel + queue
 *   block SpGEMM    kami_bsr_spgemm_*          symbolic (count/scan/fill) + numeric
 *
 * Conventions shared by every entry point:
 *   - Matrices are row-major. A is M x K. B is N x K ("B transposed", the layout of
 *     the reference's mma.row.col kernels) unless KAMI_B_KN is selected, in which
 *     case B is K x N. C is M x N. Leading dimensions are in elements, 0 = packed.
 *   - Every pointer argument that names matrix data is a device pointer. The library
 *     never frees caller memory and never allocates in kami_block_gemm.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream). Every
 *     launch is asynchronous on that stream; the only call that waits for the device is
 *     kami_bsr_spgemm_symbolic, which reads the output block count back to the host.
 *   - Errors are returned as kami_status; the library prints nothing.
 *
 * Reference interfaces replaced (paths relative to the KAMI source tree):
 *   kami_block_gemm   <- block_gemm_{1d,2d,3d}_{double,half}_mma kernels + their
 *                        launch in main(): src/block_gemm/KAMI/1d/fp64/
 *                        block_gemm_1d_double_mma.cu:55 and :263, 2d/.../:61 and :281,
 *                        3d/.../:62 and :294 (+ host slab sum :324-333), fp16 variants
 *                        src/block_gemm/KAMI/1d/fp16/block_gemm_1d_half_mma.cu:61.
 *   kami_batched_*    <- batched_gemm_{1d,2d,3d}_double_mma: src/batched_gemm/KAMI/1d/
 *                        batched_gemm_1d_double_mma.cu:59 and its launch :273.
 *   kami_bsr_spgemm_* <- spgemm_symbolic_kernel + thrust::exclusive_scan +
 *                        spgemm_symbolic_col_kernel + block_spgemm_*_half_mma:
 *                        src/sparse/spgemm/2d/block_spgemm_2d_half_mma.cu:58, :806,
 *                        :88, :171 and the host sequence :797-825.
 */
#ifndef KAMI_H
#define KAMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KAMI_VERSION_MAJOR 0
#define KAMI_VERSION_MINOR 1

typedef enum kami_status {
    KAMI_OK = 0,
    KAMI_EINVAL = 1,  /* bad argument: null pointer, bad shape, stride, alignment */
    KAMI_ENOTSUP = 2, /* valid request with no compiled instance */
    KAMI_ENOMEM = 3,  /* host or device allocation failed */
    KAMI_EHIP = 4,    /* a HIP runtime call failed, see kami_last_hip_error() */
    KAMI_ENOSPC = 5,  /* caller workspace too small */
    KAMI_ENODEV = 6   /* no gfx950 device is current */
} kami_status;

typedef enum kami_precision {
    KAMI_PREC_F64 = 0,            /* A,B,C double; native FP64 MFMA */
    KAMI_PREC_F16_F32 = 1,        /* A,B IEEE binary16; C float; FP32 accumulation */
    KAMI_PREC_BF16_F32 = 2,       /* A,B bfloat16; C float; FP32 accumulation */
    KAMI_PREC_F64_SPLIT_F16 = 3,  /* A,B,C double; FP16 slices on MFMA, FP64 recombination */
    KAMI_PREC_F64_SPLIT_BF16 = 4, /* A,B,C double; BF16 slices on MFMA, FP64 recombination */
    KAMI_PREC_F64_SPLIT_I8 = 5    /* A,B,C double; int8 digits of 62-bit fixed point on int8
                                     MFMA with exact int32 sums, FP64 recombination */
} kami_precision;

typedef enum kami_partition {
    KAMI_PART_AUTO = 0, /* tuned choice */
    KAMI_PART_1D = 1,   /* waves split M; B broadcast through LDS; A register-resident */
    KAMI_PART_2D = 2,   /* waves on an M x N grid; A and B broadcast along rows/columns */
    KAMI_PART_3D = 3    /* waves on an M x N x K grid; K partial sums reduced in LDS */
} kami_partition;

typedef enum kami_b_layout {
    KAMI_B_NK = 0, /* B stored N x K row-major (reference layout, C = A * B^T of storage) */
    KAMI_B_KN = 1  /* B stored K x N row-major */
} kami_b_layout;

/* Descriptor of one block-GEMM shape. Fill with kami_gemm_desc_init() then edit. */
typedef struct kami_gemm_desc {
    int32_t m, n, k;
    int32_t precision; /* kami_precision */
    int32_t partition; /* kami_partition */
    int32_t ranks;     /* wave-grid edge P (1D: P waves, 2D: P*P, 3D: P*P*P); 0 = default */
    int32_t b_layout;  /* kami_b_layout */
    int32_t iters;     /* >= 1: the product is accumulated `iters` times before the
                          epilogue, as the reference's NUM_ITER repeat does; 1 = plain GEMM.
                          Split precision takes iters > 1 at 32^3 and 64^3 (FP16/BF16) and
                          64^3 and 128^3 (int8): digits made once, the slice products repeated */
    int64_t lda, ldb, ldc; /* elements; 0 = packed.  Every entry point refuses (KAMI_EINVAL)
                              an operand whose rows x leading dimension, or a batch whose
                              count x stride, spans more than 2^48 bytes */
    double alpha, beta;    /* C = alpha * (A B) + beta * C */
    int32_t slices;        /* split precision: slices per operand, 0 = default */
    int32_t pipeline;      /* 0 = tuned choice; 1 = one workgroup per problem;
                              2 = persistent workgroups prefetching several K stages
                              ahead across problems (2D, B_NK, iters = 1 only) */
} kami_gemm_desc;

/* The instance a descriptor resolves to (filled by kami_block_gemm_query). */
typedef struct kami_gemm_config {
    int32_t partition;      /* resolved kami_partition */
    int32_t ranks;          /* resolved P */
    int32_t waves_m, waves_n, waves_k; /* wave grid */
    int32_t threads;        /* workgroup size */
    int32_t k_stage;        /* K elements staged through LDS per stage */
    int32_t lds_bytes;      /* static LDS of the instance */
    int32_t slices;         /* split precision slices (0 otherwise) */
    int32_t instance;       /* internal instance id */
    int32_t pipeline;       /* resolved pipeline (1 or 2) */
    int32_t prefetch;       /* K stages loaded ahead of the MFMAs */
} kami_gemm_config;

/* One problem of a batched plan. Pointers are device pointers that must stay valid
 * for the plan's lifetime. */
typedef struct kami_problem {
    int32_t m, n, k, reserved;
    int64_t lda, ldb, ldc; /* elements; 0 = packed */
    const void* A;
    const void* B;
    void* C;
} kami_problem;

typedef struct kami_batched_opts {
    int32_t b_layout;  /* kami_b_layout */
    int32_t autotune;  /* 1 = time the candidate tile variants on the device at plan
                          creation; 0 = use the built-in priors */
    int32_t use_graph; /* 1 = execute through a captured hipGraph */
    int32_t reserved;
    double alpha, beta;
} kami_batched_opts;

typedef struct kami_batched_plan kami_batched_plan;

typedef struct kami_plan_info {
    int64_t problems;
    int64_t work_items;     /* queue entries (output tiles) of one execution */
    int32_t buckets;        /* size buckets the autotuner decided */
    int32_t workgroups;     /* persistent grid size */
    double flops;           /* 2*sum(M*N*K) of one execution */
    int32_t variant_of_bucket[16]; /* chosen tile variant per bucket (first 16) */
} kami_plan_info;

/* Block-sparse row matrix: mb x nb blocks of bs x bs, blocks row-major. */
typedef struct kami_bsr {
    int32_t mb, nb, bs, nnzb;
    const int32_t* rowptr; /* device, mb + 1 */
    const int32_t* colidx; /* device, nnzb, ascending within each block row */
    const void* vals;      /* device, nnzb * bs * bs elements */
} kami_bsr;

/* ---------------------------------------------------------------- status, device */
const char* kami_status_string(kami_status s);
int32_t kami_last_hip_error(void);       /* hipError_t of the last failing HIP call on this thread */
const char* kami_version(void);
kami_status kami_device_check(void);     /* KAMI_OK when the current device is gfx950 */

/* ---------------------------------------------------------------- block GEMM */
kami_status kami_gemm_desc_init(kami_gemm_desc* d, int32_t m, int32_t n, int32_t k,
                                int32_t precision);
/* Validate a descriptor and report the instance it resolves to. Touches no device. */
kami_status kami_block_gemm_query(const kami_gemm_desc* d, kami_gemm_config* out);
/* count independent problems; problem p uses A + p*stride_a, B + p*stride_b,
 * C + p*stride_c (elements; 0 = packed: m*k, n*k, m*n). One problem per workgroup.
 * KAMI_EINVAL when count > 1 and a nonzero stride_c is below the span of one C
 * (the problems' outputs would overlap); KAMI_ENOTSUP when count workgroups of the
 * chosen instance would reach 2^32 work-items. */
kami_status kami_block_gemm(const kami_gemm_desc* d, const void* A, const void* B, void* C,
                            int64_t count, int64_t stride_a, int64_t stride_b,
                            int64_t stride_c, void* stream);

/* Times every compiled (partition, P) instance of the descriptor's shape on the given
 * buffers (C is overwritten; beta is ignored while timing) and makes the fastest the
 * KAMI_PART_AUTO choice for that shape in this process.  No reference counterpart:
 * the reference picks the partition offline in its sweep scripts
 * (scripts/fp64_block_square_GH200.sh:11-25). */
kami_status kami_block_gemm_autotune(const kami_gemm_desc* d, const void* A, const void* B,
                                     void* C, int64_t count, void* stream,
                                     kami_gemm_config* chosen);

/* FP64 GEMM through FP16 / BF16 slices at sizes beyond the block instances: same
 * arguments as kami_block_gemm with precision KAMI_PREC_F64_SPLIT_F16 or _BF16, B stored
 * N x K, iters = 1, partition AUTO or 2D; M and N multiples of 128 (or of 64), K a
 * multiple of 32.  Each problem is cut into 128 x 128 (64 x 64) tiles of C; persistent
 * workgroups walk the tiles of all problems.  Error bound as for the block form, with
 * digit exponents per row and 32-wide chunk of K.  No counterpart in the reference (it
 * has no split-precision path); it widens the split-precision block GEMM to any size. */
kami_status kami_split_gemm(const kami_gemm_desc* d, const void* A, const void* B, void* C,
                            int64_t count, int64_t stride_a, int64_t stride_b,
                            int64_t stride_c, void* stream);

/* ---------------------------------------------------------------- batched GEMM */
/* precision: KAMI_PREC_F64 or KAMI_PREC_F16_F32 or KAMI_PREC_BF16_F32.
 * Builds the work queue, autotunes per size bucket (runs kernels on `stream` when
 * opts->autotune is set: C buffers are written during tuning). */
kami_status kami_batched_plan_create(const kami_problem* probs, int64_t count,
                                     int32_t precision, const kami_batched_opts* opts,
                                     void* stream, kami_batched_plan** out);
/* One persistent launch (through the plan's hipGraph when use_graph is set).
 * A plan owns one device queue counter that every execution resets and then draws
 * from, so a plan allows one execution in flight: serialise its executions on one
 * stream (a second stream, or two replays of a captured graph running at once, would
 * race on the counter and skip or repeat tiles).  Distinct plans run concurrently. */
kami_status kami_batched_execute(kami_batched_plan* plan, void* stream);
kami_status kami_batched_plan_info(const kami_batched_plan* plan, kami_plan_info* out);
kami_status kami_batched_plan_destroy(kami_batched_plan* plan);

/* One batched GEMM without a plan and without allocating: no counterpart in the reference;
 * for callers that cannot allocate memory (code under stream capture, serving loops).  The
 * problems, precision and options mean what they mean for kami_batched_plan_create and
 * the result equals that of a plan created with the same arguments and executed once;
 * autotune and use_graph must be 0 (KAMI_EINVAL otherwise: both need memory of their
 * own), tile variants come from the tuning cache or the priors.  The work queue and its
 * counter live in a fixed device array of 2 MiB per device (held until the process ends):
 * at most 32768 queue entries (output tiles of 64 x 64, 32 x 32 or 128 x 64 by size
 * bucket), KAMI_ENOSPC beyond, checked before any device call.  The entries are built on
 * the host in static arrays and copied to the device on `stream`; the call allocates no
 * device or host memory.  Every call on a device shares that device's array: calls that
 * overlap on one device (two streams, or two host threads whose launches interleave) race
 * on the queue, so serialise them. */
kami_status kami_batched_gemm_static(const kami_problem* probs, int64_t count, int32_t precision,
                                     const kami_batched_opts* opts, void* stream);

/* ---------------------------------------------------------------- block SpGEMM */
/* Workspace for symbolic/fill of C = A * B, A with mb block rows, B with nb block cols:
 * one column bitmap per block row of C, mb * ceil(nb / 32) * 4 bytes. */
kami_status kami_bsr_spgemm_workspace_size(int32_t mb, int32_t nb, size_t* bytes);
/* rowptrC (mb+1, device) receives the exclusive scan of per-row output block counts;
 * *nnzbC receives rowptrC[mb] (read back on the host: this call waits until the work on
 * `stream` up to and including its own has stored the total).
 * KAMI_ENOTSUP when C would hold 2^31 - 1 blocks or more (int32 row pointers).
 * ws (device) is optional: with at least kami_bsr_spgemm_workspace_size bytes the rows'
 * column bitmaps are left there for kami_bsr_spgemm_fill, which then only expands them;
 * pass the same ws, unchanged, to the fill of the same A and B.  The total is written by
 * the device into a pinned host word that the host waits on (a spin of up to 200 us, then
 * a blocking wait on the stream); the first call on a host thread allocates that word
 * (64 bytes, kept until the process ends). */
kami_status kami_bsr_spgemm_symbolic(const kami_bsr* A, const kami_bsr* B, int32_t* rowptrC,
                                     int64_t* nnzbC, void* ws, size_t ws_bytes, void* stream);
/* colidxC (nnzbC, device) receives the ascending output block columns of every row: from
 * the bitmaps the symbolic call left in ws when ws holds the workspace size, else
 * rebuilt from A and B. */
kami_status kami_bsr_spgemm_fill(const kami_bsr* A, const kami_bsr* B, const int32_t* rowptrC,
                                 int32_t* colidxC, void* ws, size_t ws_bytes, void* stream);
/* valC (nnzbC * bs * bs, device) = alpha * (A * B) on the structure of rowptrC/colidxC.
 * precision: KAMI_PREC_F16_F32 or KAMI_PREC_BF16_F32 (FP16/BF16 values, float C) or
 * KAMI_PREC_F64 (double values and C). */
kami_status kami_bsr_spgemm_numeric(const kami_bsr* A, const kami_bsr* B, const int32_t* rowptrC,
                                    const int32_t* colidxC, int64_t nnzbC, void* valC,
                                    int32_t precision, double alpha, void* stream);
/* As kami_bsr_spgemm_numeric with the block products repeated `iters` >= 1 times on
 * operands held in registers, accumulated into valC = alpha * iters * (A * B): the
 * reference programs' in-kernel NUM_ITER repeat (the loop at src/sparse/spgemm/1d/
 * block_spgemm_1d_half_mma.cu:192, counted in the FLOP total at :512-516), for measuring
 * the arithmetic without the memory traffic.  iters = 1 is kami_bsr_spgemm_numeric. */
kami_status kami_bsr_spgemm_numeric_repeat(const kami_bsr* A, const kami_bsr* B,
                                           const int32_t* rowptrC, const int32_t* colidxC,
                                           int64_t nnzbC, void* valC, int32_t precision,
                                           double alpha, int32_t iters, void* stream);

/* ---------------------------------------------------------------- tuning cache */
/* kami_block_gemm_autotune and batched plans created with autotune record their choices
 * per (precision, M, N, K) and per (precision, size bucket).  Save writes them to a text
 * file; load merges a file (all or nothing: KAMI_EINVAL on any malformed record or on a
 * block choice that names no compiled instance); clear forgets them.  Replaces the
 * per-size tables of the reference's sweep scripts (scripts/fp64_block_square_GH200.sh:11-25,
 * scripts/fp64_batched_GH200.sh:96-123). */
kami_status kami_tune_save(const char* path);
kami_status kami_tune_load(const char* path);
kami_status kami_tune_clear(void);

/* ---------------------------------------------------------------- block SpMM */
/* C (mb*bs x n, row-major, ldc) = alpha * A * op(B) + beta * C, A in BSR (K = nb*bs).
 * B dense: KAMI_B_NK -> B stored n x K row-major (ldb >= K), C(i,j) = sum_k A(i,k) B[j*ldb+k],
 * the reference's layout; KAMI_B_KN -> B stored K x n (ldb >= n).  ldb / ldc of 0 mean
 * packed.  n must be a multiple of 16; bs 16 or 32.  precision: KAMI_PREC_F16_F32 /
 * KAMI_PREC_BF16_F32 (FP16/BF16 A and B, float C) or KAMI_PREC_F64.  KAMI_ENOTSUP when
 * mb * ceil(n / 64) reaches 2^24 (one workgroup per block row and 64 columns).
 * Replaces block_spmm_{1d,2d,3d}_half_mma (src/sparse/spmm/2d/block_spmm_2d_half_mma.cu:72)
 * and its host-side re-tiling bsr2tilebsr (:264-291). */
kami_status kami_bsr_spmm(const kami_bsr* A, const void* B, int32_t n, int64_t ldb,
                          int32_t b_layout, void* C, int64_t ldc, int32_t precision,
                          double alpha, double beta, void* stream);

/* ---------------------------------------------------------------- slicing, generators */
/* The digit slicing of the split-precision path, standalone: X (rows x k FP64, row-major,
 * ldx >= k, 0 = packed) is written as
 *     X(i, j) ~= 2^(e(i, c) - BETA) * sum_{s < slices} d_s(i, j) 2^(-BETA s),   c = j / 32,
 * with e(i, c) the exponent of the largest |X(i, j)| of the 32-wide chunk c of row i
 * (every |X| of the chunk below 2^e; 0 for an all-zero chunk) and digits by round to
 * nearest even, |d_0| <= 2^BETA and |d_s| <= 2^(BETA-1) after.  precision
 * KAMI_PREC_F64_SPLIT_F16 (BETA = 9, binary16 digits) or KAMI_PREC_F64_SPLIT_BF16
 * (BETA = 8, bfloat16 digits); 1 <= slices <= 8.  digits (device): `slices` planes of
 * rows x k 16-bit values, plane s at digits + s * rows * k; exps (device): int32
 * rows x ceil(k / 32).  The error per element is at most 2^(e - BETA * slices - 1);
 * finite inputs.  No counterpart in the reference (it has no split-precision path): the
 * slicing step of kami_block_gemm / kami_split_gemm, for callers that run the slice
 * products with a GEMM of their own. */
kami_status kami_split_slices(const double* X, int64_t rows, int64_t k, int64_t ldx,
                              int32_t precision, int32_t slices, uint16_t* digits,
                              int32_t* exps, void* stream);

typedef enum kami_generator {
    KAMI_GEN_RAMP = 0,   /* x[i] = i + offset */
    KAMI_GEN_RAND_R3 = 1 /* x[i] = rand_r(&seed) % 3 with seed = (unsigned)(i + offset) */
} kami_generator;

/* Fills dst (device, count values of the precision's input type: double for
 * KAMI_PREC_F64, binary16 or bfloat16 bits for KAMI_PREC_F16_F32 / _BF16_F32, rounded to
 * nearest even from the float of the value) with a known-answer input of the reference's
 * programs, generated on the device instead of on the host and copied: the ramp of
 * src/block_gemm/KAMI/1d/fp64/block_gemm_1d_double_mma.cu:224-236 and the per-element
 * rand_r(i) % 3 of src/batched_gemm/KAMI/1d/batched_gemm_1d_double_mma.cu:231-245. */
kami_status kami_generate(int32_t kind, int32_t precision, void* dst, int64_t count,
                          int64_t offset, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* KAMI_H */
