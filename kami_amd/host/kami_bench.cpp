// kami_bench: command-line driver over the C ABI (include/kami.h).
//
// It plays the role of the reference's per-kernel host programs -- build the inputs with
// the reference's generators, run the kernel, time it with events, print the
// "[hemeng_log]" line and validate against a CPU loop:
//   block   : src/block_gemm/KAMI/1d/fp64/block_gemm_1d_double_mma.cu:213-320 (FP64, ramp
//             inputs) and src/block_gemm/KAMI/1d/fp16/block_gemm_1d_half_mma.cu:205-290
//             (FP16, A = 1, B = rand() % 3); NUM_ITER is --iters
//   batched : src/batched_gemm/KAMI/1d/batched_gemm_1d_double_mma.cu:220-340 (rand_r(i) % 3
//             per element, one problem per workgroup, uniform sizes)
//   spgemm  : src/sparse/spgemm/2d/block_spgemm_2d_half_mma.cu:720-870 (srand(0) BSR
//             generator for A and B, symbolic + numeric, FLOP = 2 nnzCub bs^3)
//   spmm    : src/sparse/spmm/2d/block_spmm_2d_half_mma.cu:293-427 (srand(0) BSR generator
//             with values rand() % 2 + 1, dense B n x K, FLOP = 2 nnz(A) n)
// Shapes, precision and partition are run-time arguments instead of -D macros.  Unlike
// the reference, a failed validation makes the exit status 1.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kami.h"

namespace {

struct Args {
    std::vector<std::string> v;
    std::string get(const char* name, const char* def) const {
        for (size_t i = 0; i + 1 < v.size(); ++i)
            if (v[i] == name) return v[i + 1];
        return def;
    }
    long geti(const char* name, long def) const {
        const std::string s = get(name, "");
        return s.empty() ? def : std::strtol(s.c_str(), nullptr, 10);
    }
    double getd(const char* name, double def) const {
        const std::string s = get(name, "");
        return s.empty() ? def : std::strtod(s.c_str(), nullptr);
    }
};

#define HIP_OK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_),     \
                         __FILE__, __LINE__);                                          \
            std::exit(2);                                                              \
        }                                                                              \
    } while (0)

#define KAMI_OK_OR_DIE(x)                                                              \
    do {                                                                               \
        kami_status s_ = (x);                                                          \
        if (s_ != KAMI_OK) {                                                           \
            std::fprintf(stderr, "kami error: %s (hip %d) at %s:%d\n",                 \
                         kami_status_string(s_), kami_last_hip_error(), __FILE__,      \
                         __LINE__);                                                    \
            std::exit(2);                                                              \
        }                                                                              \
    } while (0)

// IEEE binary16 from double, round to nearest even (exact for the small integers the
// generators produce; written out so the host side needs no half type).
uint16_t to_half(double x) {
    uint16_t sign = std::signbit(x) ? 0x8000 : 0;
    double a = std::fabs(x);
    if (a == 0) return sign;
    if (a >= 65520.0) return sign | 0x7c00;
    int e;
    std::frexp(a, &e);  // a = f 2^e, f in [0.5, 1)
    int exp = e - 1;    // a = 1.m 2^exp
    if (exp < -14) {    // subnormal: units of 2^-24
        const double q = std::nearbyint(std::ldexp(a, 24));
        return sign | (uint16_t)q;
    }
    double q = std::nearbyint(std::ldexp(a, 10 - exp));  // 11-bit significand
    if (q >= 2048.0) {
        q /= 2;
        ++exp;
    }
    if (exp > 15) return sign | 0x7c00;
    return sign | (uint16_t)((exp + 15) << 10) | (uint16_t)((int)q - 1024);
}

// bfloat16 from double through float, round to nearest even.
uint16_t to_bf16(double x) {
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fff + ((u >> 16) & 1);
    return (uint16_t)(u >> 16);
}

int parse_prec(const std::string& s) {
    if (s == "f64") return KAMI_PREC_F64;
    if (s == "f16") return KAMI_PREC_F16_F32;
    if (s == "bf16") return KAMI_PREC_BF16_F32;
    if (s == "split16") return KAMI_PREC_F64_SPLIT_F16;
    if (s == "splitbf16") return KAMI_PREC_F64_SPLIT_BF16;
    std::fprintf(stderr, "unknown precision %s\n", s.c_str());
    std::exit(2);
}

int parse_part(const std::string& s) {
    if (s == "auto") return KAMI_PART_AUTO;
    if (s == "1d") return KAMI_PART_1D;
    if (s == "2d") return KAMI_PART_2D;
    if (s == "3d") return KAMI_PART_3D;
    std::fprintf(stderr, "unknown partition %s\n", s.c_str());
    std::exit(2);
}

bool half_input(int prec) { return prec == KAMI_PREC_F16_F32 || prec == KAMI_PREC_BF16_F32; }

// Device copy of `x` in the input type of `prec`.
void* upload(const std::vector<double>& x, int prec) {
    void* d = nullptr;
    if (half_input(prec)) {
        std::vector<uint16_t> h(x.size());
        for (size_t i = 0; i < x.size(); ++i)
            h[i] = prec == KAMI_PREC_F16_F32 ? to_half(x[i]) : to_bf16(x[i]);
        HIP_OK(hipMalloc(&d, h.size() * 2));
        HIP_OK(hipMemcpy(d, h.data(), h.size() * 2, hipMemcpyHostToDevice));
    } else {
        HIP_OK(hipMalloc(&d, x.size() * 8));
        HIP_OK(hipMemcpy(d, x.data(), x.size() * 8, hipMemcpyHostToDevice));
    }
    return d;
}

std::vector<double> download_c(const void* d, size_t n, int prec) {
    std::vector<double> out(n);
    if (half_input(prec)) {
        std::vector<float> h(n);
        HIP_OK(hipMemcpy(h.data(), d, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) out[i] = h[i];
    } else {
        HIP_OK(hipMemcpy(out.data(), d, n * 8, hipMemcpyDeviceToHost));
    }
    return out;
}

void print_device() {
    int dev = 0;
    hipDeviceProp_t p;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipGetDeviceProperties(&p, dev));
    std::printf("GPU %d Model: %s (%s)\n", p.pciDeviceID, p.name, p.gcnArchName);
}

const char* dim_name(int part) {
    return part == KAMI_PART_1D ? "1d" : part == KAMI_PART_2D ? "2d" : "3d";
}

// C = A B^T per problem (B stored N x K), the reference's check loop.
void cpu_gemm(const double* A, const double* B, double* C, int m, int n, int k) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0;
            for (int q = 0; q < k; ++q) s += A[(size_t)i * k + q] * B[(size_t)j * k + q];
            C[(size_t)i * n + j] = s;
        }
}

int run_block(const Args& a) {
    const int prec = parse_prec(a.get("--prec", "f64"));
    const int m = (int)a.geti("--m", 64), n = (int)a.geti("--n", m), k = (int)a.geti("--k", m);
    const long count = a.geti("--count", 1024);
    const int iters = (int)a.geti("--iters", 1);
    kami_gemm_desc d;
    KAMI_OK_OR_DIE(kami_gemm_desc_init(&d, m, n, k, prec));
    d.partition = parse_part(a.get("--part", "auto"));
    d.ranks = (int)a.geti("--ranks", 0);
    d.iters = iters;
    d.pipeline = (int)a.geti("--pipeline", 0);
    kami_gemm_config cfg;
    KAMI_OK_OR_DIE(kami_block_gemm_query(&d, &cfg));
    print_device();

    // one problem's inputs, replicated to every problem as the reference's workgroups all
    // read the same matrices: FP64 ramp A[i] = B[i] = i; FP16/BF16 A = 1, B = rand() % 3
    std::vector<double> A1((size_t)m * k), B1((size_t)n * k);
    if (half_input(prec)) {
        for (auto& x : A1) x = 1.0;
        for (auto& x : B1) x = (double)(std::rand() % 3);
    } else {
        for (size_t i = 0; i < A1.size(); ++i) A1[i] = (double)i;
        for (size_t i = 0; i < B1.size(); ++i) B1[i] = (double)i;
    }
    std::vector<double> A(A1.size() * count), B(B1.size() * count);
    for (long p = 0; p < count; ++p) {
        std::memcpy(&A[A1.size() * p], A1.data(), A1.size() * 8);
        std::memcpy(&B[B1.size() * p], B1.data(), B1.size() * 8);
    }
    void* dA = upload(A, prec);
    void* dB = upload(B, prec);
    void* dC = nullptr;
    const size_t csz = (size_t)m * n * count;
    const size_t eo = half_input(prec) ? 4 : 8;
    HIP_OK(hipMalloc(&dC, csz * eo));
    HIP_OK(hipMemset(dC, 0, csz * eo));
    HIP_OK(hipDeviceSynchronize());

    std::printf("Launching kernel with %ld blocks and %d threads per block and %d bytes of "
                "shared memory\n", count, cfg.threads, cfg.lds_bytes);
    KAMI_OK_OR_DIE(kami_block_gemm(&d, dA, dB, dC, count, 0, 0, 0, nullptr));  // warm-up
    hipEvent_t t0, t1;
    HIP_OK(hipEventCreate(&t0));
    HIP_OK(hipEventCreate(&t1));
    HIP_OK(hipEventRecord(t0, nullptr));
    KAMI_OK_OR_DIE(kami_block_gemm(&d, dA, dB, dC, count, 0, 0, 0, nullptr));
    HIP_OK(hipEventRecord(t1, nullptr));
    HIP_OK(hipEventSynchronize(t1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, t0, t1));
    const double ops = 2.0 * m * n * k * (double)count * iters;
    const double gflops = ops / (ms * 1e6), tflops = gflops / 1e3;
    std::printf("Execution time: %g ms\n", ms);
    std::printf("Performance: %g GFLOPS (%g TFLOPS)\n", gflops, tflops);
    std::printf("[hemeng_log],%s,%d,%d,%d,%d,%d,%g,%d\n", dim_name(cfg.partition), m, n, k,
                cfg.ranks, cfg.waves_m * cfg.waves_n * cfg.waves_k, tflops, cfg.threads);

    std::vector<double> C = download_c(dC, csz, prec);
    std::vector<double> R((size_t)m * n);
    cpu_gemm(A1.data(), B1.data(), R.data(), m, n, k);
    int error = 0;
    for (long p = 0; p < count && !error; ++p)
        for (size_t i = 0; i < R.size(); ++i) {
            const double got = C[R.size() * p + i] / iters;
            if (std::fabs(got - R[i]) > 1e-6) {
                std::printf("Error: problem %ld element %zu: %.17g vs %.17g\n", p, i, got, R[i]);
                error = 1;
                break;
            }
        }
    if (!error) std::printf("Validation successful!\n");
    HIP_OK(hipFree(dA));
    HIP_OK(hipFree(dB));
    HIP_OK(hipFree(dC));
    return error;
}

int run_batched(const Args& a) {
    const int prec = parse_prec(a.get("--prec", "f64"));
    if (prec != KAMI_PREC_F64 && prec != KAMI_PREC_F16_F32) {
        std::fprintf(stderr, "batched: --prec f64 or f16\n");
        return 2;
    }
    const int m = (int)a.geti("--m", 64), n = (int)a.geti("--n", m), k = (int)a.geti("--k", m);
    const long count = a.geti("--count", 1000);
    const int iters = (int)a.geti("--iters", 1);
    const int graph = (int)a.geti("--graph", 1);
    print_device();
    // rand_r(i) % 3 per element index i over the whole batch
    std::vector<double> A((size_t)m * k * count), B((size_t)n * k * count);
    for (size_t i = 0; i < A.size(); ++i) {
        unsigned s = (unsigned)i;
        A[i] = (double)(rand_r(&s) % 3);
    }
    for (size_t i = 0; i < B.size(); ++i) {
        unsigned s = (unsigned)i;
        B[i] = (double)(rand_r(&s) % 3);
    }
    void* dA = upload(A, prec);
    void* dB = upload(B, prec);
    const size_t ei = half_input(prec) ? 2 : 8, eo = half_input(prec) ? 4 : 8;
    const size_t csz = (size_t)m * n * count;
    void* dC = nullptr;
    HIP_OK(hipMalloc(&dC, csz * eo));
    HIP_OK(hipMemset(dC, 0, csz * eo));
    std::vector<kami_problem> probs(count);
    for (long p = 0; p < count; ++p) {
        kami_problem& q = probs[p];
        q.m = m;
        q.n = n;
        q.k = k;
        q.reserved = 0;
        q.lda = k;
        q.ldb = k;
        q.ldc = n;
        q.A = (const char*)dA + (size_t)p * m * k * ei;
        q.B = (const char*)dB + (size_t)p * n * k * ei;
        q.C = (char*)dC + (size_t)p * m * n * eo;
    }
    kami_batched_opts o;
    std::memset(&o, 0, sizeof(o));
    o.b_layout = KAMI_B_NK;
    o.autotune = (int)a.geti("--autotune", 1);
    o.use_graph = graph;
    o.alpha = 1.0;
    o.beta = 0.0;
    kami_batched_plan* plan = nullptr;
    KAMI_OK_OR_DIE(kami_batched_plan_create(probs.data(), count, prec, &o, nullptr, &plan));
    kami_plan_info info;
    KAMI_OK_OR_DIE(kami_batched_plan_info(plan, &info));
    std::printf("Launching kernel with %d blocks and %d threads per block (%lld work items)\n",
                info.workgroups, 256, (long long)info.work_items);
    KAMI_OK_OR_DIE(kami_batched_execute(plan, nullptr));  // warm-up
    HIP_OK(hipDeviceSynchronize());
    hipEvent_t t0, t1;
    HIP_OK(hipEventCreate(&t0));
    HIP_OK(hipEventCreate(&t1));
    HIP_OK(hipEventRecord(t0, nullptr));
    for (int it = 0; it < iters; ++it) KAMI_OK_OR_DIE(kami_batched_execute(plan, nullptr));
    HIP_OK(hipEventRecord(t1, nullptr));
    HIP_OK(hipEventSynchronize(t1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, t0, t1));
    const double ops = 2.0 * m * n * k * (double)count * iters;
    const double gflops = ops / (ms * 1e6), tflops = gflops / 1e3;
    std::printf("Execution time: %g ms\n", ms);
    std::printf("Performance: %g GFLOPS (%g TFLOPS)\n", gflops, tflops);
    std::printf("[hemeng_log],1d,%d,%d,%d,%ld,%g,%d\n", m, n, k, count, tflops, 256);

    std::vector<double> C = download_c(dC, csz, prec);
    std::vector<double> R((size_t)m * n);
    int error = 0;
    for (long p = 0; p < count && !error; ++p) {
        cpu_gemm(&A[(size_t)p * m * k], &B[(size_t)p * n * k], R.data(), m, n, k);
        for (size_t i = 0; i < R.size(); ++i)
            if (std::fabs(C[R.size() * p + i] - R[i]) > 1e-6) {
                std::printf("Error: problem %ld element %zu: %.17g vs %.17g\n", p, i,
                            C[R.size() * p + i], R[i]);
                error = 1;
                break;
            }
    }
    if (!error) std::printf("Validation successful!\n");
    KAMI_OK_OR_DIE(kami_batched_plan_destroy(plan));
    HIP_OK(hipFree(dA));
    HIP_OK(hipFree(dB));
    HIP_OK(hipFree(dC));
    return error;
}

// The reference's BSR generator: srand(0); a block is kept when rand()/RAND_MAX >
// sparsity; its values are rand() % 3 + 1, drawn right after the keep decision.
int gen_bsr(int rows, int cols, int bs, float sparsity, std::vector<int32_t>& rp,
            std::vector<int32_t>& ci, std::vector<double>& val) {
    std::srand(0);
    rp.assign(rows + 1, 0);
    ci.clear();
    val.clear();
    for (int i = 0; i < rows; ++i) {
        for (int j = 0; j < cols; ++j) {
            const float r = (float)std::rand() / (float)RAND_MAX;
            if (r > sparsity) {
                ci.push_back(j);
                for (int q = 0; q < bs * bs; ++q) val.push_back((double)(std::rand() % 3) + 1.0);
            }
        }
        rp[i + 1] = (int32_t)ci.size();
    }
    return (int)ci.size();
}

int run_spgemm(const Args& a) {
    const int prec = parse_prec(a.get("--prec", "f16"));
    if (prec == KAMI_PREC_F64_SPLIT_F16 || prec == KAMI_PREC_F64_SPLIT_BF16) {
        std::fprintf(stderr, "spgemm: --prec f16, bf16 or f64\n");
        return 2;
    }
    const int nbr = (int)a.geti("--blocks", 8), bs = (int)a.geti("--bs", 16);
    const float sparsity = (float)a.getd("--sparsity", 0.5);
    // NUM_ITER: the numeric phase repeats its block products in the kernel
    const int iters = (int)a.geti("--iters", 1);
    // the reference's measurement runs the same small product in each of its 16384 CTAs
    // (BLOCKS_PER_GRID, src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:43); here that
    // is one block-diagonal matrix of `copies` copies, each copy's product independent
    const int copies = (int)a.geti("--copies", 1);
    if (copies < 1 || (long)copies * nbr > (1L << 24)) {
        std::fprintf(stderr, "spgemm: --copies from 1 to 2^24 / blocks\n");
        return 2;
    }
    print_device();
    std::vector<int32_t> rp1, ci1;
    std::vector<double> val1;
    const int nnzb1 = gen_bsr(nbr, nbr, bs, sparsity, rp1, ci1, val1);  // A = B, as srand(0) twice
    std::printf("nnzA = %d, nnzB = %d (blocks)\n", nnzb1, nnzb1);
    const int mb = nbr * copies;
    const long nnzb_long = (long)nnzb1 * copies;
    if (nnzb_long >= (1L << 31)) {
        std::fprintf(stderr, "spgemm: too many blocks\n");
        return 2;
    }
    const int nnzb = (int)nnzb_long;
    std::vector<int32_t> rp(mb + 1), ci((size_t)nnzb);
    std::vector<double> val((size_t)nnzb * bs * bs);
    for (int c = 0; c < copies; ++c) {
        for (int i = 0; i < nbr; ++i) rp[(size_t)c * nbr + i] = c * nnzb1 + rp1[i];
        for (int q = 0; q < nnzb1; ++q) ci[(size_t)c * nnzb1 + q] = c * nbr + ci1[q];
        std::copy(val1.begin(), val1.end(), val.begin() + (size_t)c * val1.size());
    }
    rp[mb] = nnzb;
    if (copies > 1) std::printf("copies = %d (block-diagonal, %d block rows)\n", copies, mb);
    int32_t *drp = nullptr, *dci = nullptr;
    HIP_OK(hipMalloc(&drp, rp.size() * 4));
    HIP_OK(hipMalloc(&dci, std::max<size_t>(ci.size(), 1) * 4));
    HIP_OK(hipMemcpy(drp, rp.data(), rp.size() * 4, hipMemcpyHostToDevice));
    if (!ci.empty()) HIP_OK(hipMemcpy(dci, ci.data(), ci.size() * 4, hipMemcpyHostToDevice));
    void* dval = upload(val.empty() ? std::vector<double>(1, 0.0) : val, prec);
    kami_bsr A{mb, mb, bs, nnzb, drp, dci, dval};

    int32_t* drpC = nullptr;
    HIP_OK(hipMalloc(&drpC, ((size_t)mb + 1) * 4));
    hipEvent_t t0, t1, t2;
    HIP_OK(hipEventCreate(&t0));
    HIP_OK(hipEventCreate(&t1));
    HIP_OK(hipEventCreate(&t2));
    int64_t nnzbC = 0;
    int32_t* dciC = nullptr;
    // the symbolic phase as the reference times it (SPG-1D:285-370): count, scan, the
    // total read back, C's column array allocated, fill; once untimed (the first call on
    // a thread allocates the pinned word the total comes back through), then `reps`
    // times, each allocation freed before the next
    const int reps = std::max(1, (int)a.geti("--sym-reps", 1));
    float sym_ms = 0;
    for (int r = -1; r < reps; ++r) {
        if (dciC) HIP_OK(hipFree(dciC));
        dciC = nullptr;
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipEventRecord(t0, nullptr));
        KAMI_OK_OR_DIE(kami_bsr_spgemm_symbolic(&A, &A, drpC, &nnzbC, nullptr, 0, nullptr));
        HIP_OK(hipMalloc(&dciC, std::max<int64_t>(nnzbC, 1) * 4));
        KAMI_OK_OR_DIE(kami_bsr_spgemm_fill(&A, &A, drpC, dciC, nullptr, 0, nullptr));
        HIP_OK(hipEventRecord(t1, nullptr));
        HIP_OK(hipEventSynchronize(t1));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, t0, t1));
        if (r >= 0) sym_ms += ms / reps;
    }
    std::printf("nnzC = %lld (blocks)\n", (long long)nnzbC);
    std::printf("Symbolic time: %g ms\n", sym_ms);
    const size_t eo = half_input(prec) ? 4 : 8;
    void* dvalC = nullptr;
    HIP_OK(hipMalloc(&dvalC, std::max<int64_t>(nnzbC, 1) * bs * bs * eo));
    KAMI_OK_OR_DIE(kami_bsr_spgemm_numeric_repeat(&A, &A, drpC, dciC, nnzbC, dvalC, prec, 1.0,
                                                  iters, nullptr));
    HIP_OK(hipEventRecord(t1, nullptr));
    KAMI_OK_OR_DIE(kami_bsr_spgemm_numeric_repeat(&A, &A, drpC, dciC, nnzbC, dvalC, prec, 1.0,
                                                  iters, nullptr));
    HIP_OK(hipEventRecord(t2, nullptr));
    HIP_OK(hipEventSynchronize(t2));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, t1, t2));
    long long nnzCub = 0;  // sum over stored blocks (i, k) of A of the blocks in B's row k
    for (int i = 0; i < nbr; ++i)
        for (int q = rp1[i]; q < rp1[i + 1]; ++q) nnzCub += rp1[ci1[q] + 1] - rp1[ci1[q]];
    // the reference's count: per matrix, times its CTAs and NUM_ITER (:512-514)
    const double ops = 2.0 * (double)nnzCub * bs * bs * bs * copies * iters;
    const double gflops = ops / (ms * 1e6), tflops = gflops / 1e3;
    std::printf("Execution time: %g ms\n", ms);
    std::printf("Performance: %g GFLOPS (%g TFLOPS)\n", gflops, tflops);
    std::printf("[hemeng_log],2d,%lld,%d,%d,%d,%d,%g,%d\n", nnzCub, nbr * bs, nbr * bs, nbr * bs,
                1, tflops, 64);

    // check: structure against a boolean product, values against a dense block product,
    // on the first and the last copy (every copy is the same product)
    std::vector<int32_t> rpC((size_t)mb + 1), ciC(nnzbC);
    HIP_OK(hipMemcpy(rpC.data(), drpC, ((size_t)mb + 1) * 4, hipMemcpyDeviceToHost));
    if (nnzbC) HIP_OK(hipMemcpy(ciC.data(), dciC, nnzbC * 4, hipMemcpyDeviceToHost));
    std::vector<double> C = download_c(dvalC, (size_t)nnzbC * bs * bs, prec);
    int error = 0;
    std::vector<double> acc((size_t)nbr * bs * bs);
    std::vector<char> hit(nbr);
    std::vector<int> check_rows;
    for (int c : {0, copies - 1}) {
        if (c == copies - 1 && c == 0 && !check_rows.empty()) break;
        for (int i = 0; i < nbr; ++i) check_rows.push_back(c * nbr + i);
    }
    for (int i : check_rows) {
        if (error) break;
        const int c0 = (i / nbr) * nbr;  // first block column of the row's copy
        std::fill(acc.begin(), acc.end(), 0.0);
        std::fill(hit.begin(), hit.end(), 0);
        for (int q = rp[i]; q < rp[i + 1]; ++q) {
            const int kb = ci[q];
            for (int r = rp[kb]; r < rp[kb + 1]; ++r) {
                const int j = ci[r] - c0;  // column within the copy
                hit[j] = 1;
                const double* a_ = &val[(size_t)q * bs * bs];
                const double* b_ = &val[(size_t)r * bs * bs];
                double* c_ = &acc[(size_t)j * bs * bs];
                for (int x = 0; x < bs; ++x)
                    for (int z = 0; z < bs; ++z)
                        for (int y = 0; y < bs; ++y) c_[x * bs + y] += a_[x * bs + z] * b_[z * bs + y];
            }
        }
        int pos = rpC[i];
        for (int j = 0; j < nbr && !error; ++j) {
            if (!hit[j]) continue;
            if (pos >= rpC[i + 1] || ciC[pos] != j + c0) {
                std::printf("Error: structure of block row %d\n", i);
                error = 1;
                break;
            }
            for (int e = 0; e < bs * bs; ++e)
                if (std::fabs(C[(size_t)pos * bs * bs + e] / iters - acc[(size_t)j * bs * bs + e]) > 1e-6) {
                    std::printf("Error: block (%d, %d) element %d\n", i, j, e);
                    error = 1;
                    break;
                }
            ++pos;
        }
        if (!error && pos != rpC[i + 1]) {
            std::printf("Error: extra blocks in block row %d\n", i);
            error = 1;
        }
    }
    if (!error) std::printf("Validation successful!\n");
    HIP_OK(hipFree(drp));
    HIP_OK(hipFree(dci));
    HIP_OK(hipFree(dval));
    HIP_OK(hipFree(drpC));
    HIP_OK(hipFree(dciC));
    HIP_OK(hipFree(dvalC));
    return error;
}

// The SpMM program (src/sparse/spmm/2d/block_spmm_2d_half_mma.cu:293-427): the BSR
// generator with values rand() % 2 + 1, then B (n x K) rand() % 2 + 1 continuing the same
// sequence; FLOP = 2 nnz(A) n with nnz(A) counted in elements.
int run_spmm(const Args& a) {
    const int prec = parse_prec(a.get("--prec", "f16"));
    if (prec == KAMI_PREC_F64_SPLIT_F16 || prec == KAMI_PREC_F64_SPLIT_BF16) {
        std::fprintf(stderr, "spmm: --prec f16, bf16 or f64\n");
        return 2;
    }
    const int nbr = (int)a.geti("--blocks", 8), bs = (int)a.geti("--bs", 16);
    const int n = (int)a.geti("--n", nbr * bs);
    const float sparsity = (float)a.getd("--sparsity", 0.5);
    print_device();
    std::srand(0);
    std::vector<int32_t> rp(nbr + 1, 0), ci;
    std::vector<double> val;
    for (int i = 0; i < nbr; ++i) {
        for (int j = 0; j < nbr; ++j) {
            const float r = (float)std::rand() / (float)RAND_MAX;
            if (r > sparsity) {
                ci.push_back(j);
                for (int q = 0; q < bs * bs; ++q) val.push_back((double)(std::rand() % 2) + 1.0);
            }
        }
        rp[i + 1] = (int32_t)ci.size();
    }
    const int64_t K = (int64_t)nbr * bs;
    std::vector<double> B((size_t)n * K);
    for (auto& x : B) x = (double)(std::rand() % 2) + 1.0;
    const int nnzb = (int)ci.size();
    int32_t *drp = nullptr, *dci = nullptr;
    HIP_OK(hipMalloc(&drp, rp.size() * 4));
    HIP_OK(hipMalloc(&dci, std::max<size_t>(ci.size(), 1) * 4));
    HIP_OK(hipMemcpy(drp, rp.data(), rp.size() * 4, hipMemcpyHostToDevice));
    if (!ci.empty()) HIP_OK(hipMemcpy(dci, ci.data(), ci.size() * 4, hipMemcpyHostToDevice));
    void* dval = upload(val.empty() ? std::vector<double>(1, 0.0) : val, prec);
    void* dB = upload(B, prec);
    const size_t eo = half_input(prec) ? 4 : 8;
    const size_t csz = (size_t)nbr * bs * n;
    void* dC = nullptr;
    HIP_OK(hipMalloc(&dC, csz * eo));
    HIP_OK(hipMemset(dC, 0, csz * eo));
    kami_bsr A{nbr, nbr, bs, nnzb, drp, dci, dval};
    std::printf("Launching kernel with %d x %d blocks and 256 threads per block\n", nbr,
                (n + 63) / 64);
    KAMI_OK_OR_DIE(kami_bsr_spmm(&A, dB, n, 0, KAMI_B_NK, dC, 0, prec, 1.0, 0.0, nullptr));
    hipEvent_t t0, t1;
    HIP_OK(hipEventCreate(&t0));
    HIP_OK(hipEventCreate(&t1));
    HIP_OK(hipEventRecord(t0, nullptr));
    KAMI_OK_OR_DIE(kami_bsr_spmm(&A, dB, n, 0, KAMI_B_NK, dC, 0, prec, 1.0, 0.0, nullptr));
    HIP_OK(hipEventRecord(t1, nullptr));
    HIP_OK(hipEventSynchronize(t1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, t0, t1));
    const double ops = 2.0 * (double)nnzb * bs * bs * n;
    const double gflops = ops / (ms * 1e6), tflops = gflops / 1e3;
    std::printf("Execution time: %g ms\n", ms);
    std::printf("Performance: %g GFLOPS (%g TFLOPS)\n", gflops, tflops);
    std::printf("[hemeng_log],2d,%d,%d,%g,%g\n", nbr * bs, 4, ms, tflops);

    std::vector<double> C = download_c(dC, csz, prec);
    std::vector<double> R(csz, 0.0);
    for (int i = 0; i < nbr; ++i)
        for (int q = rp[i]; q < rp[i + 1]; ++q)
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c) {
                    const double x = val[(size_t)q * bs * bs + r * bs + c];
                    for (int j = 0; j < n; ++j)
                        R[((size_t)i * bs + r) * n + j] += x * B[(size_t)j * K + (size_t)ci[q] * bs + c];
                }
    int error = 0;
    for (size_t e = 0; e < csz; ++e)
        if (std::fabs(C[e] - R[e]) > 1e-6) {
            std::printf("Error: %zu %.17g %.17g\n", e, C[e], R[e]);
            error = 1;
            break;
        }
    if (!error) std::printf("Validation successful!\n");
    HIP_OK(hipFree(drp));
    HIP_OK(hipFree(dci));
    HIP_OK(hipFree(dval));
    HIP_OK(hipFree(dB));
    HIP_OK(hipFree(dC));
    return error;
}

void usage() {
    std::printf(
        "usage: kami_bench block   [--prec f64|f16|bf16|split16|splitbf16] [--m M --n N --k K]\n"
        "                          [--part auto|1d|2d|3d] [--ranks P] [--count C] [--iters I]\n"
        "                          [--pipeline 0|1|2]\n"
        "       kami_bench batched [--prec f64|f16] [--m M --n N --k K] [--count C]\n"
        "                          [--iters I] [--graph 0|1] [--autotune 0|1]\n"
        "       kami_bench spgemm  [--prec f16|bf16|f64] [--blocks B] [--bs 16|32]\n"
        "                          [--sparsity S] [--iters I] [--copies N] [--sym-reps R]\n"
        "       kami_bench spmm    [--prec f16|bf16|f64] [--blocks B] [--bs 16|32] [--n N]\n"
        "                          [--sparsity S]\n");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2 || !std::strcmp(argv[1], "--help") || !std::strcmp(argv[1], "-h")) {
        usage();
        return argc < 2 ? 2 : 0;
    }
    Args a;
    for (int i = 2; i < argc; ++i) a.v.push_back(argv[i]);
    const std::string cmd = argv[1];
    if (cmd == "block") return run_block(a);
    if (cmd == "batched") return run_batched(a);
    if (cmd == "spgemm") return run_spgemm(a);
    if (cmd == "spmm") return run_spmm(a);
    usage();
    return 2;
}
