// Batched GEMM over problems of uneven size: one persistent launch, a device work queue,
// per-size-bucket tile variants chosen by an autotuner, optional hipGraph execution.
//
// Reference: batched_gemm_{1d,2d,3d}_double_mma (src/batched_gemm/KAMI/1d/
// batched_gemm_1d_double_mma.cu:59-218) runs one uniform-size problem per CTA at fixed
// strides (:71-74) with the partition picked per size by the sweep script
// (scripts/fp64_batched_GH200.sh:96-123).  Here:
//   * problems carry their own {m, n, k, lda, ldb, ldc, A, B, C} (the kami_problem array);
//   * every problem is cut into output tiles; a tile is one queue entry
//     {problem, row0, col0, variant};
//   * a persistent grid (CUs x resident workgroups) pops entries with one atomic add
//     per tile, in descending-cost order (longest first), until the queue is empty;
//   * the variant of a tile is the tile shape and wave partition of its size bucket,
//     picked by timing candidates on the device at plan creation (or from priors);
//   * one execution is [reset the queue counter] -> [persistent kernel], captured once
//     into a hipGraph and replayed.  The counter is cleared by a one-thread kernel,
//     never by a memset node.
// M, N and K must be multiples of 16.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "kami_internal.hpp"

namespace kami {

struct WorkItem {
    int32_t prob, row0, col0, variant;
};

// Tile geometry: TM x TN output tile, WM x WN waves (4 waves), K staged KC at a time.
template <class AT, int TM_, int TN_, int WM_, int WN_, int KC_>
struct TileGeom {
    static constexpr int TM = TM_, TN = TN_, WM = WM_, WN = WN_, KC = KC_;
    static constexpr int NT = 256;
    static_assert(WM * WN * 64 == NT, "four waves");
    static constexpr int SUBM = TM / WM, SUBN = TN / WN, IM = SUBM / 16, JN = SUBN / 16;
    static constexpr int ELT = (int)sizeof(typename AT::in_t), VE = 16 / ELT;
    static constexpr int ROWB = KC * ELT + 16;
    static constexpr int A_LDS = TM * ROWB, B_LDS = TN * ROWB, STAGE = A_LDS + B_LDS;
    static constexpr int LDS = 2 * STAGE;
    static constexpr int CA = TM * KC / VE, CB = TN * KC / VE;
    static constexpr int NA = (CA + NT - 1) / NT, NB = (CB + NT - 1) / NT;
    static_assert(SUBM % 16 == 0 && SUBN % 16 == 0 && KC % AT::KSTEP == 0, "tile");
};

template <class AT, class TG>
__device__ void tile_gemm(const kami_problem& P, int row0, int col0, unsigned char* smem,
                          double alpha_d, double beta_d, bool b_kn) {
    using in_t = typename AT::in_t;
    using acc_t = typename AT::acc_t;
    using out_t = typename AT::out_t;
    using frag_t = typename AT::frag_t;
    using accv_t = typename AT::accv_t;
    constexpr int FB = AT::FRAG_BYTES;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / TG::WN, wn = wave % TG::WN;
    const int M = P.m, N = P.n, K = P.k;
    const int64_t lda = P.lda, ldb = P.ldb, ldc = P.ldc;
    const in_t* A = reinterpret_cast<const in_t*>(P.A);
    const in_t* B = reinterpret_cast<const in_t*>(P.B);
    out_t* C = reinterpret_cast<out_t*>(P.C);
    const int mlim = M - row0, nlim = N - col0;  // rows / cols of the tile inside C

    accv_t acc[TG::IM][TG::JN];
#pragma unroll
    for (int i = 0; i < TG::IM; ++i)
#pragma unroll
        for (int j = 0; j < TG::JN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[i][j][v] = 0;

    u4_t stA[TG::NA], stB[TG::NB];
    auto load_stage = [&](int k0) {
        constexpr int CPR = TG::KC / TG::VE;
#pragma unroll
        for (int i = 0; i < TG::NA; ++i) {
            const int c = tid + i * TG::NT;
            const int r = c / CPR, kc = (c % CPR) * TG::VE;
            u4_t v = {0, 0, 0, 0};
            if (c < TG::CA && r < mlim && k0 + kc < K)
                v = load_raw<16>(A + (int64_t)(row0 + r) * lda + k0 + kc);
            stA[i] = v;
        }
        if (!b_kn) {
#pragma unroll
            for (int i = 0; i < TG::NB; ++i) {
                const int c = tid + i * TG::NT;
                const int r = c / CPR, kc = (c % CPR) * TG::VE;
                u4_t v = {0, 0, 0, 0};
                if (c < TG::CB && r < nlim && k0 + kc < K)
                    v = load_raw<16>(B + (int64_t)(col0 + r) * ldb + k0 + kc);
                stB[i] = v;
            }
        } else {
            constexpr int CPRN = TG::TN / TG::VE;
#pragma unroll
            for (int i = 0; i < TG::NB; ++i) {
                const int c = tid + i * TG::NT;
                const int kr = c / CPRN, jc = (c % CPRN) * TG::VE;
                u4_t v = {0, 0, 0, 0};
                if (c < TG::CB && jc < nlim && k0 + kr < K)
                    v = load_raw<16>(B + (int64_t)(k0 + kr) * ldb + col0 + jc);
                stB[i] = v;
            }
        }
    };
    auto write_stage = [&](int buf) {
        unsigned char* base = smem + buf * TG::STAGE;
        constexpr int CPR = TG::KC / TG::VE;
#pragma unroll
        for (int i = 0; i < TG::NA; ++i) {
            const int c = tid + i * TG::NT;
            if (c < TG::CA) store_raw<16>(base + (c / CPR) * TG::ROWB + (c % CPR) * 16, stA[i]);
        }
        unsigned char* bb = base + TG::A_LDS;
        if (!b_kn) {
#pragma unroll
            for (int i = 0; i < TG::NB; ++i) {
                const int c = tid + i * TG::NT;
                if (c < TG::CB) store_raw<16>(bb + (c / CPR) * TG::ROWB + (c % CPR) * 16, stB[i]);
            }
        } else {
            constexpr int CPRN = TG::TN / TG::VE;
#pragma unroll
            for (int i = 0; i < TG::NB; ++i) {
                const int c = tid + i * TG::NT;
                if (c < TG::CB) {
                    const int kr = c / CPRN, jc = (c % CPRN) * TG::VE;
                    const in_t* e = reinterpret_cast<const in_t*>(&stB[i]);
#pragma unroll
                    for (int q = 0; q < TG::VE; ++q)
                        *reinterpret_cast<in_t*>(bb + (jc + q) * TG::ROWB + kr * TG::ELT) = e[q];
                }
            }
        }
    };
    auto compute = [&](int buf) {
        const unsigned char* base = smem + buf * TG::STAGE;
        const unsigned char* bb = base + TG::A_LDS;
#pragma unroll
        for (int ks = 0; ks < TG::KC / AT::KSTEP; ++ks) {
            const int kk = (ks * AT::KSTEP + AT::KV * (lane >> 4)) * TG::ELT;
            frag_t af[TG::IM];
#pragma unroll
            for (int im = 0; im < TG::IM; ++im)
                af[im] = bitcast<frag_t>(
                    load_raw<FB>(base + (wm * TG::SUBM + im * 16 + (lane & 15)) * TG::ROWB + kk));
#pragma unroll
            for (int jn = 0; jn < TG::JN; ++jn) {
                if (wn * TG::SUBN + jn * 16 >= nlim) continue;  // atom outside C: skip
                const frag_t bf = bitcast<frag_t>(
                    load_raw<FB>(bb + (wn * TG::SUBN + jn * 16 + (lane & 15)) * TG::ROWB + kk));
#pragma unroll
                for (int im = 0; im < TG::IM; ++im)
                    if (wm * TG::SUBM + im * 16 < mlim) AT::mma(acc[im][jn], af[im], bf);
            }
        }
    };

    const int nch = (K + TG::KC - 1) / TG::KC;
    load_stage(0);
    write_stage(0);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
        const bool has_next = c + 1 < nch;
        if (has_next) load_stage((c + 1) * TG::KC);
        compute(c & 1);
        if (has_next) write_stage((c & 1) ^ 1);
        __syncthreads();
    }

    const acc_t alpha = (acc_t)alpha_d, beta = (acc_t)beta_d;
#pragma unroll
    for (int im = 0; im < TG::IM; ++im)
#pragma unroll
        for (int jn = 0; jn < TG::JN; ++jn) {
            const int rb = wm * TG::SUBM + im * 16, cb = wn * TG::SUBN + jn * 16;
            if (rb >= mlim || cb >= nlim) continue;
            const int col = col0 + cb + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = row0 + rb + AT::out_row(lane, v);
                out_t* dst = C + (int64_t)row * ldc + col;
                acc_t val = alpha * acc[im][jn][v];
                if (beta != (acc_t)0) val += beta * (acc_t)(*dst);
                *dst = (out_t)val;
            }
        }
}

// Tile variants.  Index = variant id stored in the work items.
//   0: 64 x 64 tile, 2 x 2 waves (2D)       1: 32 x 32 tile, 2 x 2 waves (2D)
//   2: 64 x 64 tile, 4 x 1 waves (1D)       3: 128 x 64 tile, 2 x 2 waves (2D)
constexpr int kVariants = 4;
static const int kVariantTM[kVariants] = {64, 32, 64, 128};
static const int kVariantTN[kVariants] = {64, 32, 64, 64};

template <class AT>
struct Variants {
    static constexpr int KC = (sizeof(typename AT::in_t) == 8) ? 16 : 64;
    using V0 = TileGeom<AT, 64, 64, 2, 2, KC>;
    using V1 = TileGeom<AT, 32, 32, 2, 2, KC>;
    using V2 = TileGeom<AT, 64, 64, 4, 1, KC>;
    using V3 = TileGeom<AT, 128, 64, 2, 2, KC>;
    static constexpr int LDS = cmax(cmax(V0::LDS, V1::LDS), cmax(V2::LDS, V3::LDS));
};

struct BatchArgs {
    const kami_problem* probs;
    const WorkItem* items;
    unsigned int* counter;
    int64_t n_items;
    double alpha, beta;
    int32_t b_kn;
};

template <class AT>
__global__ void __launch_bounds__(256) batched_persistent_kernel(BatchArgs args) {
    using VS = Variants<AT>;
    __shared__ __attribute__((aligned(16))) unsigned char smem[VS::LDS];
    __shared__ unsigned int s_item;
    const bool b_kn = args.b_kn != 0;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(args.counter, 1u);
        __syncthreads();
        const int64_t it = s_item;
        __syncthreads();
        if (it >= args.n_items) break;  // every wave of the workgroup sees the same value
        const WorkItem w = args.items[it];
        const kami_problem P = args.probs[w.prob];
        switch (w.variant) {
            case 0: tile_gemm<AT, typename VS::V0>(P, w.row0, w.col0, smem, args.alpha, args.beta, b_kn); break;
            case 1: tile_gemm<AT, typename VS::V1>(P, w.row0, w.col0, smem, args.alpha, args.beta, b_kn); break;
            case 2: tile_gemm<AT, typename VS::V2>(P, w.row0, w.col0, smem, args.alpha, args.beta, b_kn); break;
            default: tile_gemm<AT, typename VS::V3>(P, w.row0, w.col0, smem, args.alpha, args.beta, b_kn); break;
        }
        __syncthreads();
    }
}

__global__ void reset_counter_kernel(unsigned int* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

static int size_class(int x) { return x <= 32 ? 0 : x <= 64 ? 1 : x <= 128 ? 2 : x <= 256 ? 3 : 4; }
static int bucket_of(const kami_problem& p) {
    return (size_class(p.m) * 5 + size_class(p.n)) * 5 + size_class(p.k);
}

// Prior variant of a bucket: the largest tile that the problem fills, following the
// reference's choice of fewer, larger warps for larger sizes
// (scripts/fp64_batched_GH200.sh:7-14).
static int prior_variant(int bucket) {
    const int cm = bucket / 25, cn = (bucket / 5) % 5;
    if (cm == 0 && cn == 0) return 1;
    if (cm >= 3 && cn >= 2) return 3;
    return 0;
}

}  // namespace kami

using namespace kami;

struct kami_batched_plan {
    int prec = 0;
    int64_t problems = 0;
    int64_t n_items = 0;
    int grid = 0;
    double flops = 0;
    int b_kn = 0;
    double alpha = 1.0, beta = 0.0;
    kami_problem* d_probs = nullptr;
    WorkItem* d_items = nullptr;
    unsigned int* d_counter = nullptr;
    bool use_graph = false;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> bucket_ids;      // non-empty buckets, ascending
    std::vector<int> bucket_variant;  // chosen variant per non-empty bucket
};

namespace kami {

static hipError_t launch_batched(int prec, const BatchArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, s, a.counter);
    if (prec == KAMI_PREC_F64)
        hipLaunchKernelGGL(batched_persistent_kernel<AtomF64>, dim3(grid), dim3(256), 0, s, a);
    else if (prec == KAMI_PREC_F16_F32)
        hipLaunchKernelGGL(batched_persistent_kernel<AtomF16x32>, dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(batched_persistent_kernel<AtomBF16x32>, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

static hipError_t persistent_grid(int prec, int* grid) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if (prec == KAMI_PREC_F64)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_persistent_kernel<AtomF64>, 256, 0);
    else if (prec == KAMI_PREC_F16_F32)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_persistent_kernel<AtomF16x32>, 256, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_persistent_kernel<AtomBF16x32>, 256, 0);
    if (e != hipSuccess) return e;
    *grid = std::max(1, cus * std::max(1, per_cu));
    return hipSuccess;
}

// Cut every problem into tiles of its bucket's variant; longest tiles first.
static void build_items(const std::vector<kami_problem>& probs, const std::vector<int>& bucket,
                        const int* variant_of_bucket, std::vector<WorkItem>& items) {
    struct Keyed {
        int64_t cost;
        WorkItem w;
    };
    std::vector<Keyed> tmp;
    for (size_t p = 0; p < probs.size(); ++p) {
        const int v = variant_of_bucket[bucket[p]];
        const int tm = kVariantTM[v], tn = kVariantTN[v];
        for (int r = 0; r < probs[p].m; r += tm)
            for (int c = 0; c < probs[p].n; c += tn) {
                const int64_t cost = (int64_t)std::min(tm, probs[p].m - r) *
                                     std::min(tn, probs[p].n - c) * probs[p].k;
                tmp.push_back({cost, WorkItem{(int32_t)p, r, c, v}});
            }
    }
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const Keyed& a, const Keyed& b) { return a.cost > b.cost; });
    items.clear();
    items.reserve(tmp.size());
    for (auto& k : tmp) items.push_back(k.w);
}

static kami_status upload(const void* host, size_t bytes, void** dev) {
    hipError_t e = hipMalloc(dev, bytes ? bytes : 16);
    if (e != hipSuccess) {
        g_last_hip = (int32_t)e;
        return KAMI_ENOMEM;
    }
    if (bytes) {
        e = hipMemcpy(*dev, host, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e);
    }
    return KAMI_OK;
}

static void free_plan(kami_batched_plan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    if (p->d_probs) (void)hipFree(p->d_probs);
    if (p->d_items) (void)hipFree(p->d_items);
    if (p->d_counter) (void)hipFree(p->d_counter);
    delete p;
}

}  // namespace kami

extern "C" {

kami_status kami_batched_plan_create(const kami_problem* probs, int64_t count, int32_t precision,
                                     const kami_batched_opts* opts, void* stream,
                                     kami_batched_plan** out) {
    if (!probs || !out || count <= 0 || count > 0x7fffffffLL) return KAMI_EINVAL;
    if (precision != KAMI_PREC_F64 && precision != KAMI_PREC_F16_F32 &&
        precision != KAMI_PREC_BF16_F32)
        return KAMI_ENOTSUP;
    kami_batched_opts o;
    std::memset(&o, 0, sizeof(o));
    o.alpha = 1.0;
    if (opts) o = *opts;
    if (o.b_layout != KAMI_B_NK && o.b_layout != KAMI_B_KN) return KAMI_EINVAL;
    if (o.autotune && o.beta != 0.0) return KAMI_EINVAL;  // tuning runs overwrite C
    const int ei = elt_in(precision), eo = elt_out(precision);
    std::vector<kami_problem> hp(probs, probs + count);
    double flops = 0;
    for (auto& p : hp) {
        if (p.m <= 0 || p.n <= 0 || p.k <= 0 || p.m % 16 || p.n % 16 || p.k % 16) return KAMI_EINVAL;
        if (!p.A || !p.B || !p.C) return KAMI_EINVAL;
        if (!p.lda) p.lda = p.k;
        if (!p.ldb) p.ldb = o.b_layout == KAMI_B_NK ? p.k : p.n;
        if (!p.ldc) p.ldc = p.n;
        if (p.lda < p.k || p.ldb < (o.b_layout == KAMI_B_NK ? p.k : p.n) || p.ldc < p.n)
            return KAMI_EINVAL;
        if ((p.lda * ei) % 16 || (p.ldb * ei) % 16 || ((uintptr_t)p.A % 16) ||
            ((uintptr_t)p.B % 16) || ((uintptr_t)p.C % eo))
            return KAMI_EINVAL;
        flops += 2.0 * p.m * p.n * (double)p.k;
    }
    kami_status st = device_check();
    if (st != KAMI_OK) return st;
    hipStream_t s = (hipStream_t)stream;

    kami_batched_plan* plan = new (std::nothrow) kami_batched_plan();
    if (!plan) return KAMI_ENOMEM;
    plan->prec = precision;
    plan->problems = count;
    plan->flops = flops;
    plan->b_kn = o.b_layout == KAMI_B_KN;
    plan->alpha = o.alpha;
    plan->beta = o.beta;
    plan->use_graph = o.use_graph != 0;

    std::vector<int> bucket(hp.size());
    int variant_of_bucket[125];
    bool used[125] = {false};
    for (size_t i = 0; i < hp.size(); ++i) {
        bucket[i] = bucket_of(hp[i]);
        used[bucket[i]] = true;
    }
    for (int b = 0; b < 125; ++b) variant_of_bucket[b] = prior_variant(b);

    hipError_t e = persistent_grid(precision, &plan->grid);
    if (e != hipSuccess) {
        free_plan(plan);
        return hip_fail(e);
    }
    st = upload(hp.data(), hp.size() * sizeof(kami_problem), (void**)&plan->d_probs);
    if (st == KAMI_OK) st = upload(nullptr, 0, (void**)&plan->d_counter);
    if (st != KAMI_OK) {
        free_plan(plan);
        return st;
    }

    if (o.autotune) {
        // time every variant on up to 256 problems of each bucket, keep the fastest
        hipEvent_t e0, e1;
        e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e != hipSuccess) {
            free_plan(plan);
            return hip_fail(e);
        }
        for (int b = 0; b < 125 && st == KAMI_OK; ++b) {
            if (!used[b]) continue;
            std::vector<kami_problem> sample;
            std::vector<int> sb;
            for (size_t i = 0; i < hp.size() && sample.size() < 256; ++i)
                if (bucket[i] == b) {
                    sample.push_back(hp[i]);
                    sb.push_back(0);
                }
            kami_problem* d_sample = nullptr;
            st = upload(sample.data(), sample.size() * sizeof(kami_problem), (void**)&d_sample);
            if (st != KAMI_OK) break;
            float best = 1e30f;
            for (int v = 0; v < kVariants; ++v) {
                int vb[1] = {v};
                std::vector<WorkItem> items;
                build_items(sample, sb, vb, items);
                WorkItem* d_it = nullptr;
                st = upload(items.data(), items.size() * sizeof(WorkItem), (void**)&d_it);
                if (st != KAMI_OK) break;
                BatchArgs a{d_sample, d_it, plan->d_counter, (int64_t)items.size(), o.alpha, 0.0,
                            plan->b_kn};
                e = launch_batched(precision, a, plan->grid, s);  // warm-up
                if (e == hipSuccess) e = hipEventRecord(e0, s);
                for (int rep = 0; rep < 3 && e == hipSuccess; ++rep)
                    e = launch_batched(precision, a, plan->grid, s);
                if (e == hipSuccess) e = hipEventRecord(e1, s);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                (void)hipFree(d_it);
                if (e != hipSuccess) {
                    st = hip_fail(e);
                    break;
                }
                if (ms < best) {
                    best = ms;
                    variant_of_bucket[b] = v;
                }
            }
            (void)hipFree(d_sample);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (st != KAMI_OK) {
            free_plan(plan);
            return st;
        }
    }

    std::vector<WorkItem> items;
    build_items(hp, bucket, variant_of_bucket, items);
    plan->n_items = (int64_t)items.size();
    st = upload(items.data(), items.size() * sizeof(WorkItem), (void**)&plan->d_items);
    if (st != KAMI_OK) {
        free_plan(plan);
        return st;
    }
    for (int b = 0; b < 125; ++b)
        if (used[b]) {
            plan->bucket_ids.push_back(b);
            plan->bucket_variant.push_back(variant_of_bucket[b]);
        }

    if (plan->use_graph) {
        e = hipStreamCreateWithFlags(&plan->cap_stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamBeginCapture(plan->cap_stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            BatchArgs a{plan->d_probs, plan->d_items, plan->d_counter, plan->n_items, plan->alpha,
                        plan->beta, plan->b_kn};
            hipError_t le = launch_batched(precision, a, plan->grid, plan->cap_stream);
            e = hipStreamEndCapture(plan->cap_stream, &plan->graph);
            if (le != hipSuccess) e = le;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            free_plan(plan);
            return hip_fail(e);
        }
    }
    *out = plan;
    return KAMI_OK;
}

kami_status kami_batched_execute(kami_batched_plan* plan, void* stream) {
    if (!plan) return KAMI_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (plan->use_graph) {
        e = hipGraphLaunch(plan->exec, s);
    } else {
        BatchArgs a{plan->d_probs, plan->d_items, plan->d_counter, plan->n_items, plan->alpha,
                    plan->beta, plan->b_kn};
        e = launch_batched(plan->prec, a, plan->grid, s);
    }
    if (e != hipSuccess) return hip_fail(e);
    return KAMI_OK;
}

kami_status kami_batched_plan_info(const kami_batched_plan* plan, kami_plan_info* out) {
    if (!plan || !out) return KAMI_EINVAL;
    std::memset(out, 0, sizeof(*out));
    out->problems = plan->problems;
    out->work_items = plan->n_items;
    out->buckets = (int32_t)plan->bucket_ids.size();
    out->workgroups = plan->grid;
    out->flops = plan->flops;
    for (size_t i = 0; i < plan->bucket_variant.size() && i < 16; ++i)
        out->variant_of_bucket[i] = plan->bucket_variant[i];
    return KAMI_OK;
}

kami_status kami_batched_plan_destroy(kami_batched_plan* plan) {
    free_plan(plan);
    return KAMI_OK;
}

}  // extern "C"
